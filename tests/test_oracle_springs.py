"""The oracle's springs, pins and active strain (oracle/bodies.py Springs, ActiveStrain; reference
implicit/actuation.py:319-476), pinned by what the mathematics fixes:

* a spring's gradient is the derivative of its energy; stretched (L0 < |d|), its diagonal and
  p^T H p are the energy's second derivatives; compressed, the clamped tangential part only
  ever raises them; the two ends' forces balance;
* an activation A makes the shape c + A (X - c) stress free (zero elastic gradient), and the
  activated rest area of each element is the area of its A-image;
* A = identity restores the mesh exactly.
"""
import numpy as np
import pytest

from diff_2d_fem_amd.geometry import activation_matrix
from oracle import terms as OT
from oracle.bodies import Springs
from oracle.pncg import OracleIntegrator
from tests.scenes import bilayer, spring_rig

PX = 0.5e-3


def _rig_state(seed=0):
    m, handles, rig = spring_rig(n_env=2)
    rig(m, 3)
    oi = OracleIntegrator(max_iter=40, eps=0.0, conv_check=0)
    s = oi.attach(m, 1e-3)
    rng = np.random.default_rng(seed)
    s.x = s.x + rng.normal(scale=0.3 * PX, size=s.x.shape)
    sp = [t for t in s.terms if isinstance(t, Springs)][0]
    return m, s, sp


def _energy(sp, s, x):
    x0, s.x = s.x, x
    E = np.zeros(s.B)
    sp.energy(s, E)
    s.x = x0
    return E.sum()


def _grad(sp, s, x):
    x0, g0, d0 = s.x, s.grad, s.diagH
    s.x, s.grad, s.diagH = x, np.zeros_like(x), np.zeros_like(x)
    sp.grad_diag(s)
    g, d = s.grad, s.diagH
    s.x, s.grad, s.diagH = x0, g0, d0
    return g, d


def test_spring_set_up_mirrors_the_model():
    m, s, sp = _rig_state()
    assert len(sp.a) == 6 and (sp.b < 0).sum() == 4          # two pins and two world springs
    assert np.allclose(sp.L0[sp.b >= 0], [m.actuation.springs[h].length for _, h, _ in [(0, 1, 2), (3, 4, 5)]])
    assert (sp.L0[[0, 3]] == 0.0).all()                      # pins


def test_spring_gradient_is_energy_derivative_and_forces_balance():
    m, s, sp = _rig_state(seed=1)
    s.h2 = 1.0
    x = s.x
    g, _ = _grad(sp, s, x)
    nodes = np.unique(np.concatenate([sp.a, sp.b[sp.b >= 0]]))
    eps = 1e-10
    for n in nodes:
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[n, c] += eps
            xm[n, c] -= eps
            fd = (_energy(sp, s, xp) - _energy(sp, s, xm)) / (2 * eps)
            assert abs(fd - g[n, c]) < 1e-6 * np.abs(g).max(), (n, c, fd, g[n, c])
    # a spring between two bodies pulls both ends equally and oppositely
    both = np.where(sp.b >= 0)[0]
    one = Springs(s, sp.a[both], sp.b[both], sp.k[both], sp.env[both])
    one.set(sp.anchor[both], sp.L0[both])
    g2, _ = _grad(one, s, x)
    assert np.abs(g2.sum(0)).max() < 1e-12 * np.abs(g2).max()


@pytest.mark.parametrize("stretched", [True, False])
def test_spring_curvature(stretched):
    m, s, sp = _rig_state(seed=2)
    s.h2 = 1.0
    x = s.x
    _, d = _grad(sp, s, x)
    n, ten, tang, ln = sp._geom(x)
    j = 1                                                     # the strut between the blocks
    one = Springs(s, sp.a[[j]], sp.b[[j]], sp.k[[j]], sp.env[[j]])
    one.set(sp.anchor[[j]], [0.5 * ln[j]] if stretched else [1.5 * ln[j]])
    _, dj = _grad(one, s, x)
    a = one.a[0]
    eps = 1e-11
    for c in range(2):
        xp, xm = x.copy(), x.copy()
        xp[a, c] += eps
        xm[a, c] -= eps
        h = (_grad(one, s, xp)[0][a, c] - _grad(one, s, xm)[0][a, c]) / (2 * eps)
        if stretched:
            assert abs(h - dj[a, c]) < 1e-5 * abs(dj[a, c])
        else:
            assert dj[a, c] >= h - 1e-6 * abs(h)
    p = np.random.default_rng(3).normal(scale=PX, size=x.shape)
    s.p, s.pHp = p, np.zeros(s.B)
    one.pHp(s)
    fd = (_grad(one, s, x + 1e-9 * p)[0] - _grad(one, s, x - 1e-9 * p)[0]) / 2e-9
    want = (p * fd).sum()
    if stretched:
        assert abs(s.pHp.sum() - want) < 1e-6 * abs(want)
    else:
        assert s.pHp.sum() >= want - 1e-6 * abs(want)


@pytest.mark.parametrize("fiber", [None, np.deg2rad(30.0)])
def test_activation_makes_shape_A_stress_free(fiber):
    m, regions = bilayer(n_env=1, fiber=fiber)
    m.actuation.regions[regions[0]].elems = np.arange(len(m.me.rest_area))   # the whole strip active
    A = activation_matrix(0.85, 1.10, fiber) if fiber is not None else activation_matrix(1.08)
    m.set_activation(regions[0], *((0.85, 1.10) if fiber is not None else (1.08,)))
    assert np.allclose(m.actuation.regions[regions[0]].A, A)
    oi = OracleIntegrator(max_iter=10, eps=0.0, conv_check=0)
    s = oi.attach(m, 1e-3)
    X = np.asarray(m.me.rest_pos, np.float64)
    c = X.mean(0)
    x = c + (X - c) @ A.T
    g, d = np.zeros_like(x), np.zeros_like(x)
    OT.elastic_grad_diag(x, s.tri, s.Bm, s.area, s.mu, s.lam, 1.0, 1, g, d)
    g0, _ = np.zeros_like(x), None
    OT.elastic_grad_diag(X + 0.01 * (X - c), s.tri, s.Bm0, s.area0, s.mu, s.lam, 1.0, 1, g0, np.zeros_like(x))
    assert np.abs(g).max() < 1e-9 * np.abs(g0).max()
    P = x[s.tri]
    area = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
                        - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    assert np.allclose(s.area, area, rtol=1e-12)


def test_activation_identity_restores_the_mesh():
    m, regions = bilayer(n_env=2)
    m.set_activation(regions[0], 0.9)
    oi = OracleIntegrator(max_iter=10, eps=0.0, conv_check=0)
    s = oi.attach(m, 1e-3)
    el = m.actuation.regions[regions[0]].elems
    assert not np.allclose(s.Bm[el], s.Bm0[el])
    m.set_activation(regions[0], 1.0)
    oi.attach(m, 1e-3)
    assert np.array_equal(s.Bm, s.Bm0) and np.array_equal(s.area, s.area0)
