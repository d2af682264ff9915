"""The reference's behavioural acceptance tests, as scenes and checks that run on any stepper:
the float64 oracle (CPU) or either GPU path. Each function cites the reference test it
re-expresses; bars are the reference's float64 bars unless stated.

A stepper is an adapter with step(model, dt), x(model), v(model), gap(model) (smallest
node-to-edge distance over dhat), lag(model) (the prescribed nodes' largest distance from their
targets) and collider_reaction_at(model, x, xn, friction) (collider 0's (2,) reaction with
positions x and step start xn).
"""
from __future__ import annotations

import numpy as np

from diff_2d_fem_amd import ImplicitIntegrator, Model

E, NU, RHO, G0 = 1.0e6, 0.40, 1100.0, 9.81
PX = 0.5e-3


# ------------------------------------------------------------------------- steppers
class GPUStepper:
    def __init__(self, path, **kw):
        self.integ = ImplicitIntegrator(path=path, precision="double", **kw)

    def step(self, m, dt):
        self.integ.simulate(m, dt)

    def x(self, m):
        return m.get_x()

    def v(self, m):
        return m._solver.get_v()

    def gap(self, m):
        return float(m._solver.min_distance().min()) / m.dhat

    def lag(self, m):
        return self.integ.stats(m)["bc_lag"]

    def collider_reaction(self, m, friction):
        return self.integ.collider_force(m, friction=friction)

    def collider_reaction_at(self, m, x, xn, friction, dt=1e-3):
        """The reaction on collider 0 (d2f_collider_force) with positions x and step start xn."""
        import torch
        s = self.integ.attach(m, dt)
        s.x.copy_(torch.as_tensor(x))
        s.xn.copy_(torch.as_tensor(xn))
        return self.integ.collider_force(m, friction=friction)[0]


class OracleStepper:
    def __init__(self, **kw):
        from oracle.pncg import OracleIntegrator
        self.integ = OracleIntegrator(**kw)

    def step(self, m, dt):
        self.integ.simulate(m, dt)

    def x(self, m):
        return self.integ.solver.x.copy()

    def v(self, m):
        return self.integ.solver.v.copy()

    def gap(self, m):
        s = self.integ.solver
        return float(s.ipc.min_distance(s).min()) / m.dhat

    def lag(self, m):
        return self.integ.stats()["bc_lag"]

    def collider_reaction_at(self, m, x, xn, friction, dt=1e-3):
        """The reaction on collider 0 from the oracle's barrier and friction gradients (energy
        gradients with h^2 = 1: the forces on the nodes, negated, are the reaction)."""
        from oracle import terms as T
        s = self.integ.attach(m, dt)
        c = m.collider_records()[0]
        g, dg = np.zeros_like(x), np.zeros_like(x)
        T.collider_grad_diag(c, x, s.bnd, s.dhat, s.kappa, 1.0, 0, g, dg)
        if friction:
            T.friction_grad_diag(c, x, xn, s.bnd, s.dhat, s.kappa, c["mu"], s.epsv, 1.0 / dt, 1.0,
                                 np.ones_like(x), g, dg)
        return g.sum(0)


def body(stepper, m, h):
    return stepper.x(m)[m.body_nodes(h)]


def body_v(stepper, m, h):
    return stepper.v(m)[m.body_nodes(h)]


def _soft(m, shape, center, **kw):
    return m.add_soft_body(np.ones(shape, np.uint8), physical_size=(shape[1] * PX, shape[0] * PX),
                           center=center, E=E, nu=NU, density=RHO, **kw)


def _rigid(m, shape, center, density=RHO):
    return m.add_affine_body(np.ones(shape, np.uint8), physical_size=(shape[1] * PX, shape[0] * PX),
                             center=center, density=density)


# ------------------------------------------------------------------------- prescribed bodies
def push(stepper, px_per_step, wall, steps=30):
    """A held pusher drives a free block through free space, or against a held wall
    (reference tests/test_ipc.py:256-277). Returns the model, handles, the pusher's lag
    behind its target, the smallest area ratio of any triangle and the smallest gap."""
    m = Model()
    m.body_contact = "ipc"
    m.body_friction = 0.3
    m.damping = 20.0
    pusher = _soft(m, (12, 4), (0.0, 0.0))
    block = _soft(m, (8, 8), (8 * PX, 0.0))                  # two pixels ahead of the pusher
    if wall:
        stop = _soft(m, (12, 4), (17 * PX, 0.0))             # a held body, three behind it
    m.finalize()
    m.set_kinematic_constraint(pusher, np.arange(m.body_nodes(pusher).size))
    if wall:
        m.set_kinematic_constraint(stop, np.arange(m.body_nodes(stop).size))
    gap = np.inf
    for k in range(1, steps + 1):
        m.set_kinematic_target(pusher, m.rigid_motion(pusher, (k * px_per_step * PX, 0.0)))
        stepper.step(m, 1e-3)
        gap = min(gap, stepper.gap(m))
    x = stepper.x(m)
    X = np.asarray(m.me.rest_pos, np.float64)
    t = np.asarray(m.me.tri)

    def area(P):
        return ((P[t[:, 1], 0] - P[t[:, 0], 0]) * (P[t[:, 2], 1] - P[t[:, 0], 1])
                - (P[t[:, 1], 1] - P[t[:, 0], 1]) * (P[t[:, 2], 0] - P[t[:, 0], 0]))
    assert np.isfinite(x).all() and gap > 0.0, f"smallest gap {gap:.2e} dhat"
    return m, pusher, block, stepper.lag(m), float((area(x) / area(X)).min())


def check_swept(stepper_factory):
    """Reference tests/test_ipc.py:280-295: the pusher reaches its target (lag < 1e-9 px) with
    nothing in its way, and keeps the block ahead, unsquashed; against a wall it stops more
    than 25 px short, with the block between it and the wall, squeezed to 0.05-0.9 of a
    triangle's rest area."""
    st = stepper_factory()
    m, pusher, block, lag, squash = push(st, 1.5, wall=False)
    front = body(st, m, pusher)[:, 0].max()
    assert lag < 1e-9 * PX, f"the pusher is {lag / PX:.2e} px short of its target in free space"
    assert body(st, m, block)[:, 0].min() > front, "the block is not ahead of the pusher"
    assert squash > 0.9, f"a block pushed through free space was squashed to {squash:.3f}"
    st = stepper_factory()
    m, pusher, block, lag, squash = push(st, 1.5, wall=True)
    xb = body(st, m, block)
    assert lag > 25 * PX, f"the pusher claims to be within {lag / PX:.1f} px of an unreachable target"
    assert xb[:, 0].max() < 15 * PX and xb[:, 0].min() > body(st, m, pusher)[:, 0].max(), \
        "the block is not between the pusher and the wall"
    assert 0.05 < squash < 0.9, f"smallest element at {squash:.3f} of its rest area"
    return lag / PX, squash


# ------------------------------------------------------------------------- no tunnelling
def throw(stepper, other, px_per_step, diagonal=False, steps=120):
    """A rigid block thrown at another body (reference tests/test_ipc.py:215-242): the smallest
    gap (in dhat), whether a node of the thrower entered a held target's box, and the relative
    change of the pair's linear momentum."""
    dt = 2.5e-4
    v0 = px_per_step * PX / dt
    m = Model()
    m.body_contact = "ipc"
    if diagonal:                             # A's lower right corner at B's upper left corner
        a = _rigid(m, (8, 8), (-6 * PX, 6 * PX))
        at = (4 * PX, -4 * PX)
    else:
        a = _rigid(m, (8, 8), (0.0, 0.0))
        at = (13 * PX, 0.8 * PX)
    b = _rigid(m, (8, 8), at) if other == "rigid" else _soft(m, (8, 8), at)
    m.finalize()
    if other == "held":
        m.set_kinematic_constraint(b, np.arange(m.body_nodes(b).size))
    m.set_body_velocity(a, velocity=(v0 / np.sqrt(2.0), -v0 / np.sqrt(2.0)) if diagonal else (v0, 0.0))
    mass = np.asarray(m.me.mass, np.float64)
    na, nb = m.body_nodes(a), m.body_nodes(b)

    def momentum():
        v = stepper.v(m) if stepper_started[0] else np.asarray(m._v0, np.float64)
        return (mass[np.r_[na, nb], None] * v[np.r_[na, nb]]).sum(0)
    stepper_started = [False]
    p0 = momentum()
    gap, inside = np.inf, False
    for _ in range(steps):
        stepper.step(m, dt)
        stepper_started[0] = True
        gap = min(gap, stepper.gap(m))
        if other == "held":                  # B does not move: its box is where it started
            xa, xb = body(stepper, m, a), m.rest_x(b)
            lo, hi = xb.min(0) + 1e-9, xb.max(0) - 1e-9
            inside |= bool(((xa > lo) & (xa < hi)).all(1).any())
    assert np.isfinite(stepper.x(m)).all(), "the state is not finite"
    return gap, inside, float(np.linalg.norm(momentum() - p0) / np.linalg.norm(p0))


# ------------------------------------------------------------------------- friction
def slide(stepper, kind, mu, th_deg, settle=100, run=300, average=200):
    """A block on a held slab under tilted gravity (reference tests/test_ipc.py:128-159): its
    mean sliding speed over the last `average` steps, the closed-form speed and the resting gap
    (in dhat) before the tilt."""
    c, dt = 200.0, 1e-3
    ce = (1.0 - np.exp(-dt * c)) / dt
    th = np.deg2rad(th_deg)
    m = Model()
    m.damping = c
    m.body_contact = "ipc"
    m.body_friction = mu
    slab = _soft(m, (6, 60), (0.0, 0.0))
    at = (-0.010, 0.0015 + 0.002 + 0.9 * PX / 8)
    block = _rigid(m, (8, 8), at) if kind == "rigid" else _soft(m, (8, 8), at)
    m.finalize()
    m.set_kinematic_constraint(slab, np.arange(m.body_nodes(slab).size))
    m.gravity = (0.0, -G0 * np.cos(th))
    for _ in range(settle):
        stepper.step(m, dt)
    gap = (body(stepper, m, block)[:, 1].min() - body(stepper, m, slab)[:, 1].max()) / m.dhat
    m.gravity = (G0 * np.sin(th), -G0 * np.cos(th))
    xs = []
    for _ in range(run):
        stepper.step(m, dt)
        xs.append(body(stepper, m, block)[:, 0].mean())
    v = np.diff(np.array(xs)[-average:]).mean() / dt
    if np.tan(th) > mu:
        expect = G0 * (np.sin(th) - mu * np.cos(th)) / ce
    else:
        expect = G0 * np.sin(th) / (mu * G0 * np.cos(th) / m.epsv + ce)
    return v, expect, gap


SHAPES = {"disc": dict(rx=0.006), "ellipse": dict(rx=0.008, ry=0.005),
          "box": dict(rx=0.007, ry=0.005, corner_radius=0.002),
          "triangle": dict(rx=0.007, corner_radius=0.002)}
MU_OBJ = 0.6


def object_scene(shape):
    """A soft block 16 x 40 mm held at its top above a collider placed by bisection at a
    clearance of half the barrier's reach from its nearest node (reference
    tests/test_objects.py:31-40, 84-98); "ground" is a half-space under the block."""
    from oracle.terms import sdf_eval                    # the level set, to place the collider
    m = Model()
    if shape == "ground":
        m.add_ground(0.0, mu=MU_OBJ)
    else:
        m.add_object(shape, center=(0.05, 0.0), mu=MU_OBJ, **SHAPES[shape])
    h = m.add_soft_body(np.ones((40, 16), np.uint8), physical_size=(0.016, 0.040),
                        center=(0.05, 0.060), E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    x0 = m.get_body_x(h)
    top = np.where(x0[:, 1] > x0[:, 1].max() - 1e-4)[0]
    m.set_kinematic_constraint(h, top)
    dhat = m.cfg.ipc_d_hat
    c = m.collider_records()[0]

    def clearance(oy):
        if shape == "ground":
            return float((x0[:, 1] - oy).min())
        return float(sdf_eval(c["kind"], 0.05, oy, c["rx"], c["ry"], c["rref"], x0)[0].min())
    lo, hi = x0[:, 1].min() - 0.05, x0[:, 1].min()
    assert clearance(lo) > 0.5 * dhat > clearance(hi)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if clearance(mid) > 0.5 * dhat else (lo, mid)
    m.set_collider_center(0, (0.0 if shape == "ground" else 0.05, lo))
    return m, h, x0


SLIP = 3e-6                                            # tangential slide of the probe step [m]


def friction_ratio(stepper_factory):
    """|F_friction| / |F_normal| of the collider reaction when the block has slid SLIP to the
    right within one step of 1 ms (reference tests/test_objects.py:99-110), per shape."""
    out = {}
    for shape in list(SHAPES) + ["ground"]:
        m, h, x0 = object_scene(shape)
        st = stepper_factory()
        x = np.asarray(m.me.rest_pos, np.float64)
        xn = x.copy()
        xn[:, 0] -= SLIP
        Fn = np.asarray(st.collider_reaction_at(m, x, xn, False), np.float64)
        Ff = np.asarray(st.collider_reaction_at(m, x, xn, True), np.float64) - Fn
        assert np.linalg.norm(Fn) > 1e-3, f"{shape}: no normal reaction, not in contact"
        out[shape] = float(np.linalg.norm(Ff) / np.linalg.norm(Fn))
    return out


# ------------------------------------------------------------------------- resting contact, grip
def stack(stepper):
    """Two soft blocks stacked on the ground (reference tests/test_ipc.py:163-181): gap between
    them (in dhat), the upper block's largest speed and the ground's normal reaction over both
    weights."""
    m = Model()
    m.body_contact = "ipc"
    m.add_ground(height=0.0, mu=0.6)
    m.damping = 60.0
    reach = PX / 8
    lower = _soft(m, (12, 16), (0.0, 6 * PX + 0.9 * reach))
    upper = _soft(m, (12, 16), (0.0, 18 * PX + 0.9 * reach + 0.5 * PX))
    m.finalize()
    m.gravity = (0.0, -G0)
    for _ in range(300):
        stepper.step(m, 2e-3)
    gap = (body(stepper, m, upper)[:, 1].min() - body(stepper, m, lower)[:, 1].max()) / m.dhat
    weight = np.asarray(m.me.mass).sum() * G0
    Fy = abs(stepper.integ.collider_force(m, friction=False)[0][1])
    return gap, float(np.abs(body_v(stepper, m, upper)).max()), Fy / weight


def pinch_and_lift(stepper, mu, lift=4.0e-3):
    """Two soft pads close on a free rigid block on the ground and lift (reference
    tests/test_ipc.py:185-204): how far the block rose over the lift."""
    m = Model()
    m.body_contact = "ipc"
    m.body_friction = mu
    m.add_ground(height=0.0, mu=0.5)
    block = _rigid(m, (16, 16), (0.0, 8 * PX + 0.9 * PX / 8), density=600.0)
    pads = [_soft(m, (16, 8), (sgn * 14 * PX, 8 * PX + 0.8e-3)) for sgn in (-1, +1)]
    m.finalize()
    m.gravity = (0.0, -G0)
    for sgn, h in zip((-1, +1), pads):                       # the pads are driven by their backs
        X = m.rest_x(h)
        back = X[:, 0].min() if sgn < 0 else X[:, 0].max()
        m.set_kinematic_constraint(h, np.where(np.abs(X[:, 0] - back) < 1e-9)[0])
    y0 = m.rest_x(block)[:, 1].mean()
    n_close, n_lift, n_hold = 100, 200, 50
    for k in range(1, n_close + n_lift + n_hold + 1):
        a = min(k / n_close, 1.0)
        b = min(max(k - n_close, 0) / n_lift, 1.0)
        for sgn, h in zip((-1, +1), pads):
            X = m.rest_x(h)
            X[:, 0] -= sgn * 2.25 * PX * a
            X[:, 1] += lift * b
            m.set_kinematic_target(h, X)
        stepper.step(m, 1e-3)
    assert np.isfinite(stepper.x(m)).all() and stepper.gap(m) > 0.0
    return (body(stepper, m, block)[:, 1].mean() - y0) / lift


# ------------------------------------------------------------------------- rigid bodies
def _com(stepper, m):
    mass = np.asarray(m.me.mass, np.float64)
    return (mass[:, None] * stepper.x(m)).sum(0) / mass.sum()


def slope(stepper, kind, th_deg, mu, settle=300, run=400, average=60):
    """A block on the ground under tilted gravity (reference tests/test_rigid.py:162-189): the
    ground's normal reaction over the weight after settling, the sliding speed over the last
    `average` steps and the closed-form speed."""
    th = np.deg2rad(th_deg)
    c, dt = 20.0, 1e-3
    m = Model()
    m.add_ground(height=0.0, mu=mu)
    m.damping = c
    reach = 0.5e-3 / 8
    shape, size, at = (8, 8), (0.004, 0.004), (0.0, 0.003 + 0.9 * reach)
    if kind == "rigid":
        m.add_affine_body(np.ones(shape, np.uint8), physical_size=size, center=at, density=RHO)
    else:
        m.add_soft_body(np.ones(shape, np.uint8), physical_size=size, center=at, E=E, nu=NU, density=RHO)
    m.finalize()
    m.gravity = (0.0, -G0 * np.cos(th))
    for _ in range(settle):
        stepper.step(m, dt)
    W = np.asarray(m.me.mass).sum() * G0 * np.cos(th)
    N = abs(stepper.integ.collider_force(m, friction=False)[0][1])
    m.gravity = (G0 * np.sin(th), -G0 * np.cos(th))
    xs = []
    for _ in range(run):
        stepper.step(m, dt)
        xs.append(_com(stepper, m)[0])
    v = np.diff(np.array(xs)[-average:]).mean() / dt
    ce = (1.0 - np.exp(-dt * c)) / dt
    if np.tan(th) > mu:
        expect = G0 * (np.sin(th) - mu * np.cos(th)) / ce
    else:
        expect = G0 * np.sin(th) / (mu * G0 * np.cos(th) / m.epsv + ce)
    return N / W, v, expect


def impact(stepper):
    """A rigid block thrown at a soft one, no damping (reference tests/test_rigid.py:252-274):
    the soft block's speed after, the relative change of linear momentum and the final gap."""
    m = Model()
    m.damping = 0.0
    sq = dict(physical_size=(0.004, 0.004))
    rigid = m.add_affine_body(np.ones((8, 8), np.uint8), center=(0.0, 0.0), density=RHO, **sq)
    soft = m.add_soft_body(np.ones((8, 8), np.uint8), center=(0.0065, 0.0), E=E, nu=NU, density=RHO, **sq)
    m.finalize()
    m.set_body_velocity(rigid, velocity=(0.3, 0.0))
    mass = np.asarray(m.me.mass, np.float64)
    p0 = (mass[:, None] * np.asarray(m._v0, np.float64)).sum(0)
    for _ in range(80):
        stepper.step(m, 2.5e-4)
    v = stepper.v(m)
    p1 = (mass[:, None] * v).sum(0)
    ns = m.body_nodes(soft)
    vs = (mass[ns, None] * v[ns]).sum(0) / mass[ns].sum()
    gap = body(stepper, m, soft)[:, 0].min() - body(stepper, m, rigid)[:, 0].max()
    return float(vs[0]), float(np.linalg.norm(p1 - p0) / np.linalg.norm(p0)), float(gap)
