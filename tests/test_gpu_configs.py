"""The benchmark configurations as parity cases: small instances of each generator
(diff_2d_fem_amd.workloads) stepped on both device paths against the float64 oracle with a
fixed iteration count.

  c1 strip_on_ground      the quickstart strip
  c2 random_blobs         random blob bodies, per-pixel stiffness, pinned bottom nodes
  c3 pads_lift_block      pads gripping a rigid block (IPC between bodies, kinematic jaws)
  c4 pneumatic_fingers    chambered fingers under pressure
  c5 star_meshes          Delaunay star meshes with rigid regions and tendons

float64: positions within 1e-9 of the body size and identical iteration counts; mixed
(scene path): within 1e-5 of the body size. Mixed precision forms the displacement gradient
H = F - I in float64 and the stress from it in float32, so the float stress keeps its relative
accuracy near the rest state.

c5 also runs at the sizes of its definition (2,000-2,500 and 18,000-20,000 triangles per
scene); the larger scene's nodes do not fit one workgroup's LDS, so the scene kernel takes its
colour-by-colour element pass there.
"""
import numpy as np
import pytest

from diff_2d_fem_amd import ImplicitIntegrator
from diff_2d_fem_amd import workloads as W
from oracle.pncg import OracleIntegrator

pytestmark = pytest.mark.gpu

CASES = {
    "c1": (lambda: W.strip_on_ground(2), 4),
    "c2": (lambda: W.random_blobs(3), 4),
    "c3": (lambda: W.pads_lift_block(2), 12),
    "c4": (lambda: W.pneumatic_fingers(2), 3),
    "c5": (lambda: W.star_meshes(2, tri_range=(300, 500)), 3),
}
KW = dict(max_iter=40, eps=0.0, conv_check=0)
_oracle = {}


def _oracle_run(name):
    if name not in _oracle:
        make, n = CASES[name]
        w = make()
        oi = OracleIntegrator(**KW)
        for k in range(1, n + 1):
            w["command"](k)
            oi.simulate(w["model"], w["dt"])
        _oracle[name] = (oi.solver.x.copy(), oi.stats()["iters"].copy(), w["model"].me.rest_pos)
    return _oracle[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path,precision", [("plain", "double"), ("scene", "double"), ("scene", "mixed")])
def test_config_parity(name, path, precision):
    make, n = CASES[name]
    w = make()
    gi = ImplicitIntegrator(path=path, precision=precision, **KW)
    for k in range(1, n + 1):
        w["command"](k)
        gi.simulate(w["model"], w["dt"])
    xo, it_o, rest = _oracle_run(name)
    x = w["model"].get_x()
    err = np.abs(x - xo).max() / w["size"]
    moved = np.abs(xo - rest).max() / w["size"]
    print(f"{name} {path} {precision}: max |dx| / size = {err:.3e} (motion {moved:.2e})")
    assert moved > 1e-5
    if precision == "double":
        assert err < 1e-9
        assert (gi.stats(w["model"])["iters"] == it_o).all()
    else:
        assert err < 1e-5


@pytest.mark.parametrize("tri_range,n_scenes", [((2000, 2500), 2), ((18000, 20000), 1)])
@pytest.mark.parametrize("path,precision", [("plain", "double"), ("scene", "double"), ("scene", "mixed")])
def test_c5_full_size(tri_range, n_scenes, path, precision):
    """Star meshes at the triangle counts of their definition, two steps of 40 iterations."""
    def run(integ):
        w = W.star_meshes(n_scenes, tri_range=tri_range)
        for k in range(1, 3):
            w["command"](k)
            integ.simulate(w["model"], w["dt"])
        return w
    key = ("c5", tri_range)
    if key not in _oracle:
        oi = OracleIntegrator(**KW)
        wo = run(oi)
        _oracle[key] = (oi.solver.x.copy(), wo["model"].me.rest_pos)
    xo, rest = _oracle[key]
    w = run(ImplicitIntegrator(path=path, precision=precision, **KW))
    T = len(w["model"].me.rest_area)
    assert tri_range[0] * n_scenes <= T <= tri_range[1] * n_scenes
    if path == "scene" and tri_range[0] >= 18000:
        assert w["model"]._solver.nacc_lds == 0          # the colour-by-colour element pass
    err = np.abs(w["model"].get_x() - xo).max() / w["size"]
    moved = np.abs(xo - rest).max() / w["size"]
    print(f"c5 {tri_range} {path} {precision}: T = {T}, max |dx| / size = {err:.3e} (motion {moved:.2e})")
    assert moved > 1e-5
    assert err < (1e-9 if precision == "double" else 1e-5)


def test_longest_first_launch_order():
    """With more scenes than compute units the scene kernel launches the scenes in order of
    decreasing size (solver.longest_first): the same steps as the launch in scene order, 1e-9 of
    the body size (float64), on C2's uneven blobs with the compute-unit count lowered so that the
    order applies to 12 scenes."""
    def run(longest_first):
        w = W.random_blobs(12)
        m = w["model"]
        integ = ImplicitIntegrator(path="scene", precision="double", max_iter=60, eps=0.0, conv_check=0)
        integ.attach(m, w["dt"])
        s = m._solver
        s.n_cu, s.longest_first = 4, longest_first
        for k in range(1, 4):
            w["command"](k)
            integ.simulate(m, w["dt"])
        return m.get_x(), s, w
    x1, s1, w = run(True)
    x0, s0, _ = run(False)
    assert s1._order_t is not None and s0._order_t is None
    order = s1._order_t.cpu().numpy()
    assert sorted(order.tolist()) == list(range(12)) and order.tolist() != list(range(12))
    assert np.abs(x1 - x0).max() <= 1e-9 * w["size"]


def test_lean_and_full_scene_kernels_agree():
    """The scene kernel's lean variant (a batch without tendons, springs, pressure, rigid modes,
    node contact or line search) and its full one give the same steps: the grip scene alone takes
    the lean kernel; with a spring of negligible stiffness added it takes the full one."""
    from tests.scenes import grip, close_jaws

    def run(spring):
        m, pads = grip(n_env=2)
        if spring:                    # slack at rest and 1e-30 N/m: a no-op term
            c = m.rest_x(pads[0][0]).mean(0)
            m.add_spring(pads[0][0], c, point_b=c + [1e-4, 0.0], stiffness=1e-30, rest_length=1e-4)
        integ = ImplicitIntegrator(path="scene", precision="mixed", max_iter=60, eps=0.0, conv_check=0)
        for k in range(1, 6):
            close_jaws(m, pads, k)
            integ.simulate(m, 1e-3)
        return m.get_x(), m._solver
    x0, s0 = run(False)
    x1, s1 = run(True)
    assert s0.scene_struct().n_sp == 0 and s1.scene_struct().n_sp == 1
    assert np.abs(x1 - x0).max() <= 1e-12
