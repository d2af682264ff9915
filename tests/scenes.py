"""Scene builders shared by the tests: small versions of the benchmark configurations."""
import numpy as np

from diff_2d_fem_amd import Model


def strip(n_env=1, E=0.5e6, nu=0.40, rho=1100.0, mu=0.8, shape=(6, 72), size=(0.054, 0.0045),
          center=(0.0, 0.0030), gravity=True):
    """The quickstart strip lying on the ground (reference README.md:38-58), n_env copies."""
    m = Model()
    m.add_ground(0.0, mu=mu)
    for i in range(n_env):
        m.add_soft_body(np.ones(shape, np.uint8), physical_size=size, center=center, E=E, nu=nu,
                        density=rho, env_id=i)
    m.finalize()
    if gravity:
        m.gravity = (0.0, -9.81)
    return m


def grip(n_env=1, px=0.5e-3, block=8, pad=(8, 4), friction=0.8):
    """Two soft pads either side of a rigid block on the ground, pads' back faces driven
    (a small version of the pinch-and-lift scene, reference examples/pinch_and_lift.py:29-66)."""
    m = Model()
    m.body_friction = friction
    m.add_ground(0.0, mu=0.5)
    pads = []
    for i in range(n_env):
        m.add_affine_body(np.ones((block, block), np.uint8), physical_size=(block * px, block * px),
                          center=(0.0, block * px / 2 + 0.9 * px / 8), density=600.0, env_id=i)
        ph, pw = pad
        pads.append([m.add_soft_body(np.ones((ph, pw), np.uint8), physical_size=(pw * px, ph * px),
                                     center=(s * (block * px / 2 + 2 * px + pw * px / 2), ph * px / 2 + 0.8e-3),
                                     E=1.0e6, nu=0.40, density=1100.0, env_id=i) for s in (-1, 1)])
    m.finalize()
    m.gravity = (0.0, -9.81)
    for pair in pads:
        for s, p in zip((-1, 1), pair):
            X = m.rest_x(p)
            back = X[:, 0].min() if s < 0 else X[:, 0].max()
            m.set_kinematic_constraint(p, np.where(np.abs(X[:, 0] - back) < 1e-9)[0])
    return m, pads


def close_jaws(m, pads, k, n_close=10, squeeze=2.25, px=0.5e-3, lift=0.0):
    a = min(k / n_close, 1.0)
    for pair in pads:
        for s, p in zip((-1, 1), pair):
            X = m.rest_x(p)
            X[:, 0] -= s * squeeze * px * a
            X[:, 1] += lift
            m.set_kinematic_target(p, X)


def cantilever():
    """The reference Newton oracle's cantilever (tests/golden/static_fixtures.npz, written by
    the reference's tests/reference_static.py): a 4 x 20 px beam, left column held, a dead
    load of 20 g on every node. Returns the model and the fixture (cant_x is the static
    equilibrium to a 1e-12 residual)."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "static_fixtures.npz"))
    m = Model()
    h = m.add_soft_body(np.ones((4, 20), np.uint8), physical_size=(0.02, 0.004), center=(0.0, 0.0),
                        E=1e5, nu=0.4, density=1000.0)
    m.finalize()
    m.set_kinematic_constraint(h, f["cant_fixed"])
    m.set_external_force(h, f["cant_fext"])
    return m, f


def pressed_block(mode):
    """The reference adjoint test's scene (tests/test_adjoint.py:35-50): a 10 x 6 px soft block,
    top row held and driven (press 3 mm, shear 2 mm over the steps) onto a disc ("disc"), a
    rounded box ("box"), or nothing within reach ("elastic"). Returns (model, handle, rest
    positions of the body, the held top nodes)."""
    m = Model()
    oy = -1.0 if mode == "elastic" else 0.0335
    if mode == "box":
        m.add_object("box", center=(0.05, oy), rx=0.006, ry=0.006, corner_radius=0.002, mu=0.0)
    else:
        m.add_disc(center=(0.05, oy), radius=0.006, mu=0.0)
    h = m.add_soft_body(np.ones((10, 6), np.uint8), physical_size=(0.012, 0.024), center=(0.05, 0.052),
                        E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    x0 = m.get_body_x(h)
    top = np.where(x0[:, 1] > x0[:, 1].max() - 1e-5)[0]
    m.set_kinematic_constraint(h, top)
    return m, h, x0, top


def press_targets(x0, top, t, T):
    x = x0.copy()
    x[top, 1] = x0[top, 1] - 0.003 * t / T
    x[top, 0] = x0[top, 0] + 0.002 * t / T
    return x


def node_stack(n_env=2, px=0.5e-3, mu=0.0, block_mu=None, gap=1.55, tilt=2.0, rigid=False):
    """A block resting on a held slab with the "node" contact between bodies (a small version of
    the sliding and slope set-ups of reference tools/measure_contact.py:23-49): the block's lower
    face starts `gap` px above the slab's upper face, just outside the contact radius of 1.5 px,
    falls into contact within a few steps and settles on the penalty, and gravity is tilted by
    `tilt` m/s^2 along the slab so that the pairs slide. Scene i shifts the block by i/4 px so
    that the scenes differ. mu: model.body_friction; block_mu: the block's
    own coefficient (per-body friction); rigid: the block is an affine body."""
    m = Model()
    m.body_contact = "node"
    m.body_friction = mu
    slabs = []
    for i in range(n_env):
        slabs.append(m.add_soft_body(np.ones((4, 20), np.uint8), physical_size=(20 * px, 4 * px),
                                     center=(0.0, -2.0 * px), E=1.0e6, nu=0.40, density=1100.0, env_id=i))
        add = m.add_affine_body if rigid else m.add_soft_body
        kw = dict(physical_size=(6 * px, 6 * px), center=(0.25 * i * px, (3.0 + gap) * px),
                  density=1100.0, env_id=i, mu=block_mu)
        if not rigid:
            kw.update(E=0.5e6, nu=0.40)
        add(np.ones((6, 6), np.uint8), **kw)
    m.finalize()
    m.gravity = (tilt, -9.81)
    for h in slabs:
        m.set_kinematic_constraint(h, np.arange(len(m.rest_x(h))))
    return m


def node_press(px=0.5e-3, gap=1.4):
    """A 6 x 10 px soft block standing on a held slab through the "node" contact, its lower face
    `gap` px above the slab's (inside the 1.5 px radius, so the same pairs act throughout), top
    row held and driven: pressed by 0.04 px and sheared by 0.1 px over the steps (see
    node_press_targets). The adjoint's scene for contact between bodies. Returns (model, block
    handle, the block's rest positions, its top nodes)."""
    m = Model()
    m.body_contact = "node"
    m.body_friction = 0.0
    slab = m.add_soft_body(np.ones((4, 20), np.uint8), physical_size=(20 * px, 4 * px),
                           center=(0.0, -2.0 * px), E=1.0e6, nu=0.40, density=1100.0)
    h = m.add_soft_body(np.ones((10, 6), np.uint8), physical_size=(6 * px, 10 * px),
                        center=(0.0, (5.0 + gap) * px), E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    m.set_kinematic_constraint(slab, np.arange(len(m.rest_x(slab))))
    x0 = m.get_body_x(h)
    top = np.where(x0[:, 1] > x0[:, 1].max() - 1e-9)[0]
    m.set_kinematic_constraint(h, top)
    return m, h, x0, top


def node_press_targets(x0, top, t, T, px=0.5e-3):
    x = x0.copy()
    x[top, 1] = x0[top, 1] - 0.04 * px * t / T
    x[top, 0] = x0[top, 0] + 0.1 * px * t / T
    return x


def ipc_press(px=0.5e-3, gap=0.5, region=False):
    """A 6 x 10 px soft block standing on a soft slab (its bottom row held) through the IPC barrier
    between bodies, frictionless: the block's lower face starts `gap` dhat above the slab's
    (dhat = px / 8, so the same node-edge terms act throughout), its top row held and driven:
    pressed by 0.2 dhat and sheared by 0.1 px over the steps (ipc_press_targets). region: rows
    3-6 of the block are one rigid (affine) region with the block's own modulus. The adjoint's
    scene for the IPC and affine tangents. Returns (model, block handle, the block's rest
    positions, its top nodes)."""
    m = Model()
    m.body_contact = "ipc"
    m.body_friction = 0.0
    slab = m.add_soft_body(np.ones((4, 20), np.uint8), physical_size=(20 * px, 4 * px),
                           center=(0.0, -2.0 * px), E=1.0e6, nu=0.40, density=1100.0)
    h = m.add_soft_body(np.ones((10, 6), np.uint8), physical_size=(6 * px, 10 * px),
                        center=(0.0, 5.0 * px + gap * px / 8), E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    if region:
        mask = np.zeros((10, 6), bool)
        mask[3:7, :] = True
        m.add_affine_region(h, mask, E=1.0e6, nu=0.40)
    Xs = m.rest_x(slab)
    m.set_kinematic_constraint(slab, np.where(Xs[:, 1] < Xs[:, 1].min() + 1e-9)[0])
    x0 = m.get_body_x(h)
    top = np.where(x0[:, 1] > x0[:, 1].max() - 1e-9)[0]
    m.set_kinematic_constraint(h, top)
    return m, h, x0, top


def ipc_press_targets(x0, top, t, T, px=0.5e-3):
    x = x0.copy()
    x[top, 1] = x0[top, 1] - 0.2 * (px / 8) * t / T
    x[top, 0] = x0[top, 0] + 0.1 * px * t / T
    return x


def spring_rig(n_env=2, px=0.5e-3):
    """Springs and pins (reference implicit/actuation.py:368-476, sim_model.py:988-1060): per
    scene a soft block pinned to the world at its top-left corner, a spring with a rest length
    from its right face to a second block, and a world spring from the second block to an
    anchor; gravity on. Scene i shifts everything by i px. Returns (model, spring handles,
    rigging function(model, step) that drives the spring lengths and the anchor)."""
    m = Model()
    pairs = []
    for i in range(n_env):
        dx = i * px
        a = m.add_soft_body(np.ones((6, 8), np.uint8), physical_size=(8 * px, 6 * px), center=(dx, 0.0),
                            E=1.0e6, nu=0.40, density=1100.0, env_id=i)
        b = m.add_soft_body(np.ones((6, 6), np.uint8), physical_size=(6 * px, 6 * px),
                            center=(dx + 12 * px, 0.0), E=0.5e6, nu=0.40, density=1100.0, env_id=i)
        pairs.append((a, b, dx))
    m.finalize()
    m.gravity = (0.0, -9.81)
    handles = []
    for a, b, dx in pairs:
        pin = m.add_pin(a, (dx - 4 * px, 3 * px), stiffness=5.0e5)
        sp = m.add_spring(a, (dx + 4 * px, 0.0), b, (dx + 9 * px, 0.0), stiffness=2.0e4)
        ws = m.add_spring(b, (dx + 15 * px, 3 * px), None, (dx + 16 * px, 5 * px), stiffness=1.0e4,
                          rest_length=1.5 * px)
        handles.append((pin, sp, ws))
    L0 = [m.actuation.springs[sp].length for _, sp, _ in handles]

    def rig(model, k):
        for (pin, sp, ws), l0 in zip(handles, L0):
            model.set_spring_length(sp, l0 * (1.0 - 0.03 * min(k, 8)))      # the strut pulls in
            model.set_spring_anchor(ws, np.asarray(model.actuation.springs[ws].anchor) + np.array([0.0, 0.05 * px]))
    return m, handles, rig


def bilayer(n_env=2, px=0.5e-3, fiber=None):
    """A 6 x 24 px cantilever, left column held, whose lower half is active material (the
    bimetal strip of reference tests/test_actuation.py:305-327). Scene i activates by
    (1 - 0.01 (i + 1)); `fiber` None: isotropic, else the fibre angle. Returns (model, region
    handles)."""
    m = Model()
    hs = []
    for i in range(n_env):
        hs.append(m.add_soft_body(np.ones((6, 24), np.uint8), physical_size=(24 * px, 6 * px),
                                  center=(0.0, 0.0), E=1.0e6, nu=0.40, density=1100.0, env_id=i))
    m.finalize()
    regions = []
    for i, h in enumerate(hs):
        X = m.rest_x(h)
        m.set_kinematic_constraint(h, np.where(X[:, 0] < X[:, 0].min() + 1e-9)[0])
        lower = np.zeros((6, 24), bool)
        lower[:3] = True
        regions.append(m.add_active_region(h, lower, fiber=fiber))
    return m, regions
