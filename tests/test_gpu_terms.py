"""GPU parity of the terms the grip and strip scenes leave out, on both device paths, against
the float64 CPU oracle with a fixed iteration count (eps = 0, no convergence poll):
pressure chambers (reference implicit/actuation.py:89-151), tendons in tension and length
mode (:169-316), external force (:37-72), the energy line search (implicit/solver.py:403-425),
disc / box / triangle colliders (core/sdf.py:45-137), kinematic bodies moved by pose, free
fall through the rigid-mode coarse space (implicit/affine.py:489-563), and per-environment
convergence with different designs per environment. Positions must agree to 1e-9 of the
body size, and iteration counts exactly.

Also the state edges of the model on the device: reset(env_ids) and set_body_x/pose/velocity
followed by a step must give what a fresh model in that state gives.
"""
import numpy as np
import pytest

from diff_2d_fem_amd import ImplicitIntegrator, Model
from oracle.pncg import OracleIntegrator

pytestmark = pytest.mark.gpu

PATHS = ["plain", "scene"]
FIXED = dict(max_iter=50, eps=0.0, conv_check=0)


def _compare(make, n_steps, dt, path, kw, size, each=None, tol=1e-9):
    mo, mg = make(), make()
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path=path, precision="double", **kw)
    for k in range(1, n_steps + 1):
        if each is not None:
            each(mo, k)
            each(mg, k)
        oi.simulate(mo, dt)
        gi.simulate(mg, dt)
    err = np.abs(mg.get_x() - oi.solver.x).max() / size
    moved = np.abs(oi.solver.x - mo.me.rest_pos).max() / size
    print(f"{path}: max |dx| / size = {err:.3e} (motion {moved:.3e} of the size), iters "
          f"{gi.stats(mg)['iters']} vs {oi.stats()['iters']}", flush=True)
    assert moved > 1e-4                      # the scene did something
    assert err < tol
    assert (gi.stats(mg)["iters"] == oi.stats()["iters"]).all()
    return oi, gi, mo, mg


def _pin_left(m, h):
    X = m.rest_x(h)
    m.set_kinematic_constraint(h, np.where(np.abs(X[:, 0] - X[:, 0].min()) < 1e-9)[0])


def finger(n_env=2):
    """A pneumatic finger: a 10 x 40 px beam with three sealed chambers, left end held, a
    different pressure per environment and ambient pressure on one."""
    mask = np.ones((10, 40), np.uint8)
    for c0 in (6, 18, 30):
        mask[3:7, c0:c0 + 5] = 0
    m = Model()
    hs = [m.add_soft_body(mask, physical_size=(0.02, 0.005), center=(0.0, 0.01), E=0.5e6, nu=0.40,
                          density=1100.0, env_id=i) for i in range(n_env)]
    m.finalize()
    for i, h in enumerate(hs):
        _pin_left(m, h)
        for c in m.add_pressure_chambers(h):
            m.set_pressure(c, 20e3 * (1 + i) * (1 + 0.2 * c))
    amb = m.add_ambient_pressure(hs[-1])
    m.set_pressure(amb, 2e3)
    return m


@pytest.mark.parametrize("path", PATHS)
def test_pressure_chambers(path):
    _compare(finger, 5, 2e-3, path, FIXED, 0.02)


def tendon_strip(n_env=2):
    """A held beam bent by a cable along its lower edge: tension-controlled in environment 0,
    length-controlled in environment 1; gravity on."""
    m = Model()
    hs = [m.add_soft_body(np.ones((5, 40), np.uint8), physical_size=(0.03, 0.00375),
                          center=(0.0, 0.02), E=0.3e6, nu=0.40, density=1100.0, env_id=i)
          for i in range(n_env)]
    m.finalize()
    m.gravity = (0.0, -9.81)
    pts = np.stack([np.linspace(-0.014, 0.0145, 8), np.full(8, 0.0185)], 1)
    for i, h in enumerate(hs):
        _pin_left(m, h)
        if i == 0:
            t = m.add_tendon(h, through=pts)
            m.set_tendon_tension(t, 30.0)
        else:
            t = m.add_tendon(h, through=pts, stiffness=2.0e3)
            m.set_tendon_length(t, 0.97 * m.tendon_length(t))
    return m


@pytest.mark.parametrize("path", PATHS)
def test_tendons(path):
    _compare(tendon_strip, 5, 2e-3, path, FIXED, 0.03)


def tendon_out_and_back(n_pts):
    """Two tendons per scene along the held beam's lower edge, the second out and back (its
    eyelets repeat nodes, and share nodes with the first): with 8 points the scene kernel's
    parallel tendon pass runs, with 40 its scratch does not fit and one thread walks them."""
    def make():
        m = Model()
        hs = [m.add_soft_body(np.ones((5, 40), np.uint8), physical_size=(0.03, 0.00375),
                              center=(0.0, 0.02), E=0.3e6, nu=0.40, density=1100.0, env_id=i) for i in range(2)]
        m.finalize()
        m.gravity = (0.0, -9.81)
        out = np.stack([np.linspace(-0.014, 0.0145, n_pts), np.full(n_pts, 0.0185)], 1)
        for i, h in enumerate(hs):
            _pin_left(m, h)
            t = m.add_tendon(h, through=out[: n_pts // 2])
            m.set_tendon_tension(t, 20.0)
            t = m.add_tendon(h, through=np.concatenate([out, out[-2::-1]]), stiffness=1.5e3)
            m.set_tendon_length(t, (0.97 - 0.02 * i) * m.tendon_length(t))
        return m
    return make


@pytest.mark.parametrize("n_pts", [8, 40])
def test_tendons_repeated_eyelets(n_pts):
    make = tendon_out_and_back(n_pts)
    ey = np.asarray(make().actuation.tendons[1].nodes)
    assert len(np.unique(ey)) < len(ey)                 # the out-and-back cable repeats nodes
    _compare(make, 5, 2e-3, "scene", FIXED, 0.03)


def pushed_block(n_env=2):
    """A soft block on the ground pushed sideways by an external force, with gravity."""
    m = Model()
    m.add_ground(0.0, mu=0.3)
    hs = [m.add_soft_body(np.ones((8, 8), np.uint8), physical_size=(0.008, 0.008),
                          center=(0.0, 0.0041), E=0.2e6, nu=0.45, density=1000.0, env_id=i)
          for i in range(n_env)]
    m.finalize()
    m.gravity = (0.0, -9.81)
    for i, h in enumerate(hs):
        n = len(m.body_nodes(h))
        m.set_external_force(h, np.tile([[4.0 * (i + 1) / n, 0.0]], (n, 1)))
    return m


@pytest.mark.parametrize("path", PATHS)
def test_external_force(path):
    _compare(pushed_block, 6, 2e-3, path, FIXED, 0.008)


@pytest.mark.parametrize("path", PATHS)
def test_line_search(path):
    """The strip dropped on the ground with the energy line search on."""
    def make():
        m = Model()
        m.add_ground(0.0, mu=0.8)
        for i in range(2):
            m.add_soft_body(np.ones((4, 30), np.uint8), physical_size=(0.03, 0.004),
                            center=(0.0, 0.0025 + 0.001 * i), E=0.5e6, nu=0.40, density=1100.0, env_id=i)
        m.finalize()
        m.gravity = (0.0, -9.81)
        return m
    kw = dict(FIXED, line_search=True, n_backtrack=4)
    _compare(make, 6, 2e-3, path, kw, 0.03)


def on_colliders(n_env=4):
    """Soft squares falling with a sideways velocity onto a disc, a box, a triangle and an
    ellipse collider (one environment above each; colliders are shared by every environment)."""
    m = Model()
    m.add_disc((-0.03, 0.0), 0.008, mu=0.4)
    m.add_object("box", (0.0, 0.0), 0.008, 0.006, corner_radius=1e-3, mu=0.6)
    m.add_object("triangle", (0.03, 0.0), 0.008, 0.008, corner_radius=1e-3, mu=0.2)
    m.add_object("ellipse", (0.06, 0.0), 0.009, 0.005, mu=0.5)
    for i in range(n_env):
        m.add_soft_body(np.ones((6, 6), np.uint8), physical_size=(0.006, 0.006),
                        center=(-0.03 + 0.03 * i + 0.001, 0.0125), E=0.3e6, nu=0.42, density=1000.0,
                        env_id=i)
    m.finalize()
    m.gravity = (0.0, -9.81)
    for h in range(n_env):
        m.set_body_velocity(h, (0.2, -0.4))
    return m


@pytest.mark.parametrize("path", PATHS)
def test_disc_box_triangle_ellipse_colliders(path):
    oi, gi, mo, mg = _compare(on_colliders, 8, 2e-3, path, FIXED, 0.006)
    assert oi.solver.clr_min.max() < 1e-3   # every body came near its collider


def kinematic_pusher(n_env=2):
    """A kinematic bar swept by pose into a soft block resting on the ground."""
    m = Model()
    m.body_friction = 0.5
    m.add_ground(0.0, mu=0.5)
    bars = []
    for i in range(n_env):
        m.add_soft_body(np.ones((6, 6), np.uint8), physical_size=(0.006, 0.006), center=(0.0, 0.0031),
                        E=0.3e6, nu=0.42, density=1000.0, env_id=i)
        bars.append(m.add_kinematic_body(np.ones((6, 2), np.uint8), physical_size=(0.002, 0.006),
                                         center=(-0.0045, 0.0035), env_id=i))
    m.finalize()
    m.gravity = (0.0, -9.81)
    m._bars = bars
    return m


@pytest.mark.parametrize("path", PATHS)
def test_kinematic_body_by_pose(path):
    def each(m, k):
        for i, h in enumerate(m._bars):
            m.set_kinematic_pose(h, translation=(min(k, 8) * 1.5e-4 * (1 + 0.5 * i), 0.0),
                                 angle=0.01 * k * i)
    _compare(kinematic_pusher, 10, 1e-3, path, FIXED, 0.006, each=each)


@pytest.mark.parametrize("path", PATHS)
def test_free_fall_through_rigid_modes(path):
    """A soft body in free fall with no contact: the rigid-mode coarse space carries it."""
    def make():
        m = Model()
        for i in range(2):
            m.add_soft_body(np.ones((6, 20), np.uint8), physical_size=(0.02, 0.006), center=(0.0, 0.1),
                            E=0.5e6 * (1 + i), nu=0.40, density=1100.0, env_id=i)
        m.finalize()
        m.gravity = (0.0, -9.81)
        m.set_body_velocity(0, (0.0, 0.0), omega=2.0)
        return m
    _compare(make, 5, 2e-3, path, dict(max_iter=25, eps=0.0, conv_check=0), 0.02)


@pytest.mark.parametrize("path", PATHS)
def test_convergence_per_environment(path):
    """Different stiffness per environment, default tolerances: each environment stops at its
    own convergence check, and the iteration counts must match the oracle's exactly."""
    def make():
        m = Model()
        m.add_ground(0.0, mu=0.8)
        for i in range(3):
            m.add_soft_body(np.ones((6, 36), np.uint8), physical_size=(0.027, 0.0045),
                            center=(0.0, 0.0026), E=0.3e6 * 3 ** i, nu=0.40, density=1100.0, env_id=i)
        m.finalize()
        m.gravity = (0.0, -9.81)
        return m
    mo, mg = make(), make()
    oi, gi = OracleIntegrator(), ImplicitIntegrator(path=path, precision="double")
    for _ in range(6):
        oi.simulate(mo, 2e-3)
        gi.simulate(mg, 2e-3)
    it_o, it_g = oi.stats()["iters"], gi.stats(mg)["iters"]
    print(f"{path}: iters {it_g} vs {it_o}")
    assert len(set(it_o.tolist())) > 1            # the environments stopped at different checks
    assert (it_g == it_o).all()
    assert np.abs(mg.get_x() - oi.solver.x).max() / 0.027 < 1e-7


def _three_strips():
    m = Model()
    m.add_ground(0.0, mu=0.8)
    for i in range(3):
        m.add_soft_body(np.ones((4, 24), np.uint8), physical_size=(0.024, 0.004), center=(0.0, 0.0025),
                        E=0.4e6, nu=0.40, density=1100.0, env_id=i)
    m.finalize()
    m.gravity = (0.0, -9.81)
    return m


@pytest.mark.parametrize("path", PATHS)
def test_reset_and_set_state(path):
    """reset(env_ids) puts the chosen environments back at rest; set_body_pose with a
    velocity moves a body; a step from there equals a step of a fresh model placed the same
    way (the environments are independent, so the untouched one is unaffected)."""
    kw = dict(max_iter=40, eps=0.0, conv_check=0)
    m = _three_strips()
    gi = ImplicitIntegrator(path=path, precision="double", **kw)
    for _ in range(4):
        gi.simulate(m, 2e-3)
    x_before = m.get_x()
    m.reset([1])
    m.set_body_pose(2, translation=(0.0, 0.002), angle=0.05, velocity=(0.1, 0.0), omega=1.0)
    env = np.asarray(m.me.node2env)
    assert np.array_equal(m.get_x()[env == 1], m.me.rest_pos[env == 1])
    assert np.array_equal(m.get_x()[env == 0], x_before[env == 0])
    gi.simulate(m, 2e-3)

    f = _three_strips()
    gf = ImplicitIntegrator(path=path, precision="double", **kw)
    gf.attach(f, 2e-3)       # the solver starts from rest, as the reference's does: place after
    f.set_body_pose(2, translation=(0.0, 0.002), angle=0.05, velocity=(0.1, 0.0), omega=1.0)
    gf.simulate(f, 2e-3)
    for e in (1, 2):
        err = np.abs(m.get_x()[env == e] - f.get_x()[env == e]).max()
        print(f"{path}: env {e} reset/set vs fresh: {err:.3e}")
        assert err < 1e-12


@pytest.mark.parametrize("path,precision,tol", [("plain", "double", 1e-8), ("scene", "double", 1e-8),
                                                ("scene", "mixed", 1e-5)])
def test_cantilever_reaches_reference_static_equilibrium(path, precision, tol):
    """The held, loaded beam stepped with a one-second step size reaches the static
    equilibrium of the reference's Newton oracle (golden fixture). Mixed precision forms the
    stress in float32 from a float64 displacement gradient; at this deflection (0.7 of the
    length: large rotation, small strain) its equilibrium lies within 1e-5 of the length."""
    from tests.scenes import cantilever
    m, f = cantilever()
    gi = ImplicitIntegrator(path=path, precision=precision, max_iter=400, eps=0.0, conv_check=0)
    for _ in range(5):
        gi.simulate(m, 1.0)
    err = np.abs(m.get_x() - f["cant_x"]).max() / 0.02
    print(f"{path} {precision}: |x - x_static| / length = {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("name", ["beam", "grip"])
def test_reference_table_scenes(name, precision):
    """Two scenes of the reference's throughput table (workloads.TABLE_SCENES, after its
    tools/benchmark_envs.py) on the scene path against the oracle: the beam bent by a tendon
    through all 81 pixel columns (the kernel's parallel tendon pass; a different tension per
    environment), and the pads gripping the rigid block (IPC with friction between bodies,
    driven pads). Float64 within 1e-9 of the body length, mixed within 1e-5."""
    from diff_2d_fem_amd.workloads import TABLE_SCENES
    mo, kw, cmd_o, dt, _ = TABLE_SCENES[name](2)
    mg, _, cmd_g, _, _ = TABLE_SCENES[name](2)
    kw = dict(kw, **FIXED)
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path="scene", precision=precision, **kw)
    for k in range(1, 4):
        cmd_o(k)
        cmd_g(k)
        oi.simulate(mo, dt)
        gi.simulate(mg, dt)
    size = np.ptp(mo.me.rest_pos[:, 0])
    err = np.abs(mg.get_x() - oi.solver.x).max() / size
    moved = np.abs(oi.solver.x - mo.me.rest_pos).max() / size
    print(f"{name} {precision}: max |dx| / size = {err:.3e} (motion {moved:.3e})", flush=True)
    assert moved > 1e-5
    assert err < (1e-9 if precision == "double" else 1e-5)
    assert (gi.stats(mg)["iters"] == oi.stats()["iters"]).all()
