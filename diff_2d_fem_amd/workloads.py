"""Seeded scene generators for the benchmark configurations.

Every random choice of scene i comes from np.random.default_rng([seed, i]), so a scene is the
same whether it is built alone or inside a batch.

  strip_on_ground   the quickstart strip (reference README.md:38-58): mask ones(6, 72),
                    0.054 x 0.0045 m, E 0.5 MPa, nu 0.40, rho 1100, ground mu 0.8, gravity,
                    dt 2 ms. One copy per scene.
  random_blobs      seeded random blob bodies: H ~ U[4, 32], W ~ U[16, 96] px of 0.75 mm, a
                    smoothed-noise blob (largest 4-connected piece, >= 8 px), per-pixel E
                    log-uniform in [1e5, 5e6] Pa, nu ~ U[0.40, 0.45], rho 1100, 2-6 pinned
                    nodes on the bottom face, ground mu 0.6, gravity, dt 2 ms.
  pads_lift_block   two soft pads 48 x 16 px (E 0.3 MPa, nu 0.45, rho 1100) on kinematic jaws
                    either side of a rigid 16 x 16 px block (rho 600) of side U[12, 20] mm
                    (pixel = side / 16) resting on the ground; IPC contact between bodies with
                    mu 0.6, ground mu 0.6, gravity, dt 1 ms. The jaws close from 2 px apart to a
                    quarter pixel past touching over 10 steps, then lift half the block side
                    over 200 steps.
  pneumatic_fingers a 12 x 96 px finger (pixel 0.25 mm, E 0.5 MPa, nu 0.4, rho 1100) with 6-10
                    sealed rectangular chambers (3-6 x 4-8 px, 1 px apart and from the border),
                    left column held, pressure U[0, 40] kPa per chamber, gravity, dt 2 ms, and a
                    random target point for the tip.
  star_meshes       a Delaunay mesh body on a random star outline r = R (1 + sum_k a_k cos(k t +
                    phi_k)), k <= 5, with a target triangle count U[2k, 20k]; 1-3 disjoint
                    round rigid regions; 2 tendons through interior waypoints, tension U[0, 50]
                    N/m; ground mu 0.6, gravity, dt 2 ms.

The scenes of the reference's throughput table (tools/benchmark_envs.py:27-77, README.md:206-210),
with the parameters of that tool, for the same protocol: TABLE_SCENES[name](n) returns
(model, integrator keywords, command, dt, lead steps).
  beam   an 8 x 80 px cantilever (0.040 x 0.004 m, E 1 MPa, nu 0.40, rho 1100, drag 50 /s), its
         left column held, bent by a tendon through every pixel column at y = 1 mm, tension
         10 + 20 i / (n - 1) N/m in environment i; no self-contact, no friction; dt 1 ms.
  strip  strip_on_ground; dt 2 ms.
  grip   a rigid 16 x 16 px block (pixel 0.5 mm, rho 600) on the ground (mu 0.5) between two
         soft 16 x 8 px pads (E 1 MPa) whose outer columns are driven: 2.25 px inwards over 40
         steps, then up 20 um per step; friction between bodies 0.8; dt 1 ms.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

from .model import Model


def strip_on_ground(n_env=1, seed=0):
    m = Model()
    m.add_ground(0.0, mu=0.8)
    for i in range(n_env):
        m.add_soft_body(np.ones((6, 72), np.uint8), physical_size=(0.054, 0.0045),
                        center=(0.0, 0.0030), E=0.5e6, nu=0.40, density=1100.0, env_id=i)
    m.finalize()
    m.gravity = (0.0, -9.81)
    return dict(model=m, dt=2e-3, command=lambda step: None, bodies=list(range(n_env)),
                size=0.054, name="strip_on_ground")


def _blob(rng):
    while True:
        H = int(rng.integers(4, 33))
        W = int(rng.integers(16, 97))
        noise = ndimage.gaussian_filter(rng.random((H, W)), sigma=1.5)
        mask = noise > np.median(noise)
        lab, n = ndimage.label(mask, structure=np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], bool))
        if n == 0:
            continue
        sizes = ndimage.sum(mask, lab, range(1, n + 1))
        keep = lab == (1 + int(np.argmax(sizes)))
        if keep.sum() >= 8:
            return keep.astype(np.uint8)


def random_blobs(n_env=1024, seed=0, px=0.75e-3):
    m = Model()
    m.add_ground(0.0, mu=0.6)
    handles, pins = [], []
    for i in range(n_env):
        rng = np.random.default_rng([seed, i])
        mask = _blob(rng)
        H, W = mask.shape
        E = np.exp(rng.uniform(np.log(1e5), np.log(5e6), size=(H, W)))
        nu = float(rng.uniform(0.40, 0.45))
        low = int(np.nonzero(mask.any(1))[0][0])            # lowest solid row
        cy = 0.5 * H * px - low * px + px                   # that row's bottom edge 1 px above ground
        h = m.add_soft_body(mask, physical_size=(W * px, H * px), center=(0.0, cy), E=E, nu=nu,
                            density=1100.0, env_id=i)
        handles.append(h)
        pins.append(rng)
    m.finalize()
    m.gravity = (0.0, -9.81)
    for h, rng in zip(handles, pins):
        X = m.rest_x(h)
        bottom = np.nonzero(np.abs(X[:, 1] - X[:, 1].min()) < 1e-9)[0]
        k = int(min(rng.integers(2, 7), len(bottom)))
        m.set_kinematic_constraint(h, np.sort(rng.choice(bottom, size=k, replace=False)))
    return dict(model=m, dt=2e-3, command=lambda step: None, bodies=handles,
                size=float(np.mean([np.ptp(m.rest_x(h), 0).max() for h in handles])),
                name="random_blobs")


def pads_lift_block(n_env=256, seed=0, n_close=10, n_lift=200, scene_ids=None):
    """Two soft pads on kinematic jaws grip a rigid block and lift it, one scene per
    environment (the protocol of reference examples/pinch_and_lift.py:23-24, 51-56, at larger
    sizes): the jaws close from 2 px apart to a quarter pixel past touching over n_close
    steps, then rise by half the block's side over n_lift steps and hold.

    scene_ids: the scene indices to build (default range(n_env)); scene i draws its block
    side from np.random.default_rng([seed, i]) and becomes the environment at its position
    in the list."""
    m = Model()
    m.body_friction = 0.6
    m.add_ground(0.0, mu=0.6)
    pads, blocks, pxs = [], [], []
    ids = list(range(n_env)) if scene_ids is None else [int(v) for v in scene_ids]
    for i, sid in enumerate(ids):
        rng = np.random.default_rng([seed, sid])
        side = float(rng.uniform(12e-3, 20e-3))
        px = side / 16.0
        pxs.append(px)
        blocks.append(m.add_affine_body(np.ones((16, 16), np.uint8), physical_size=(side, side),
                                        center=(0.0, 0.5 * side + 0.9 * px / 8.0), density=600.0,
                                        env_id=i))
        gap = 2.0 * px
        pads.append([m.add_soft_body(np.ones((48, 16), np.uint8), physical_size=(16 * px, 48 * px),
                                     center=(s * (0.5 * side + gap + 8.0 * px), 24.0 * px + px),
                                     E=3.0e5, nu=0.45, density=1100.0, env_id=i) for s in (-1, 1)])
    m.finalize()
    m.gravity = (0.0, -9.81)
    # the pads' driven back faces, all scenes at once: global ids, rest positions, side (-1
    # left, +1 right) and scene of every driven node (the step reads no other node's target)
    nodes, X0, side, env = [], [], [], []
    for i, pair in enumerate(pads):
        for s, p in zip((-1, 1), pair):
            X = m.rest_x(p)
            back = X[:, 0].min() if s < 0 else X[:, 0].max()
            driven = np.where(np.abs(X[:, 0] - back) < 1e-9)[0]
            m.set_kinematic_constraint(p, driven)
            nodes.append(m.body_nodes(p)[driven])
            X0.append(X[driven])
            side.append(np.full(len(driven), float(s)))
            env.append(np.full(len(driven), i))
    nodes, X0, side, env = (np.concatenate(a) for a in (nodes, X0, side, env))
    close = np.array([2.25 * px for px in pxs])             # to a quarter pixel past touching
    lift = np.array([8.0 * px for px in pxs])

    def command(step):
        c = close * min(step / n_close, 1.0)
        up = lift * min(max(step - n_close, 0) / n_lift, 1.0)
        Y = X0.copy()
        Y[:, 0] -= side * c[env]
        Y[:, 1] += up[env]
        m.set_node_targets(nodes, Y)
    return dict(model=m, dt=1e-3, command=command, bodies=blocks, pads=pads,
                size=float(np.mean([16 * px for px in pxs])),
                sides=np.array([16 * px for px in pxs]), name="pads_lift_block")


def _chambers(rng, H=12, W=96):
    """6-10 non-overlapping rectangles (3-6 x 4-8 px), each kept 1 px from the border and from
    the others, as a list of (row, col, h, w)."""
    want = int(rng.integers(6, 11))
    taken = np.zeros((H, W), bool)                 # chamber pixels grown by one
    rects = []
    for _ in range(2000):
        if len(rects) == want:
            break
        h, w = int(rng.integers(3, 7)), int(rng.integers(4, 9))
        r, c = int(rng.integers(1, H - h)), int(rng.integers(1, W - w))
        if r + h > H - 1 or c + w > W - 1 or taken[r - 1:r + h + 1, c - 1:c + w + 1].any():
            continue
        taken[r - 1:r + h + 1, c - 1:c + w + 1] = True
        rects.append((r, c, h, w))
    return rects


def pneumatic_fingers(n_env=512, seed=0, px=0.25e-3):
    m = Model()
    hs, pressures, targets = [], [], []
    for i in range(n_env):
        rng = np.random.default_rng([seed, i])
        mask = np.ones((12, 96), np.uint8)
        for r, c, h, w in _chambers(rng):
            mask[r:r + h, c:c + w] = 0
        hs.append(m.add_soft_body(mask, physical_size=(96 * px, 12 * px), center=(0.0, 0.02),
                                  E=0.5e6, nu=0.4, density=1100.0, env_id=i))
        pressures.append(rng)
    m.finalize()
    m.gravity = (0.0, -9.81)
    chambers = []
    for i, (h, rng) in enumerate(zip(hs, pressures)):
        X = m.rest_x(h)
        m.set_kinematic_constraint(h, np.nonzero(np.abs(X[:, 0] - X[:, 0].min()) < 1e-9)[0])
        ch = m.add_pressure_chambers(h)
        for c in ch:
            m.set_pressure(c, float(rng.uniform(0.0, 40e3)))
        chambers.append(ch)
        tip = X[np.argmax(X[:, 0])]
        targets.append(tip + rng.uniform(-1.0, 1.0, 2) * 4e-3)
    return dict(model=m, dt=2e-3, command=lambda step: None, bodies=hs, chambers=chambers,
                targets=np.asarray(targets), size=96 * px, name="pneumatic_fingers")


def _star_mesh(rng, R, n_tri):
    """A Delaunay mesh of a random star outline with about n_tri triangles: boundary points
    every h along the outline, a jittered triangular lattice inside, triangles kept when
    their centroid is inside; vertices and triangles sorted for locality."""
    from scipy.spatial import Delaunay
    K = 5
    a = rng.uniform(-0.08, 0.08, K) / np.arange(1, K + 1)
    ph = rng.uniform(0.0, 2 * np.pi, K)

    def radius(t):
        return R * (1.0 + sum(a[k] * np.cos((k + 1) * t + ph[k]) for k in range(K)))
    area = 0.5 * np.mean(radius(np.linspace(0.0, 2 * np.pi, 4096, endpoint=False)) ** 2) * 2 * np.pi
    h = np.sqrt(2.0 * area / (np.sqrt(3.0) / 2.0 * n_tri))     # two triangles per lattice point
    t = np.linspace(0.0, 2 * np.pi, 8192, endpoint=False)
    P = np.stack([radius(t) * np.cos(t), radius(t) * np.sin(t)], 1)
    seg = np.linalg.norm(np.roll(P, -1, 0) - P, axis=1)
    arc = np.concatenate([[0.0], np.cumsum(seg)])
    nb = max(12, int(round(arc[-1] / h)))
    tb = np.interp(np.linspace(0.0, arc[-1], nb, endpoint=False), arc, np.concatenate([t, [2 * np.pi]]))
    B = np.stack([radius(tb) * np.cos(tb), radius(tb) * np.sin(tb)], 1)
    xs = np.arange(-1.3 * R, 1.3 * R, h)
    ys = np.arange(-1.3 * R, 1.3 * R, h * np.sqrt(3.0) / 2.0)
    gx, gy = np.meshgrid(xs, ys)
    gx = gx + 0.5 * h * (np.arange(len(ys))[:, None] % 2)
    G = np.stack([gx.ravel(), gy.ravel()], 1) + rng.uniform(-0.1, 0.1, (gx.size, 2)) * h
    rg, tg = np.linalg.norm(G, axis=1), np.arctan2(G[:, 1], G[:, 0])
    G = G[rg < radius(tg) - 0.7 * h]
    V = np.concatenate([B, G])
    T = Delaunay(V).simplices
    c = V[T].mean(1)
    T = T[np.linalg.norm(c, axis=1) < radius(np.arctan2(c[:, 1], c[:, 0]))]
    used = np.unique(T)
    remap = np.full(len(V), -1)
    order = used[np.lexsort((V[used, 1], np.floor(V[used, 0] / (8 * h))))]
    remap[order] = np.arange(len(order))
    V, T = V[order], remap[T]
    e1, e2 = V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]]
    flip = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0] < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    T = T[np.argsort(T.min(1), kind="stable")]
    return V, T, h


def star_meshes(n_env=4096, seed=0, R=10e-3, tri_range=(2000, 20000), progress=None):
    """`progress(done, n_env, what)`, when given, is called after every 256 meshes and after the
    model is finalized (building the full batch takes minutes)."""
    m = Model()
    m.add_ground(0.0, mu=0.6)
    hs, extra = [], []
    for i in range(n_env):
        rng = np.random.default_rng([seed, i])
        V, T, h = _star_mesh(rng, R, int(rng.integers(tri_range[0], tri_range[1] + 1)))
        V = V + np.array([0.0, -V[:, 1].min() + 2.0 * h])
        hs.append(m.add_soft_mesh(V, T, E=float(np.exp(rng.uniform(np.log(2e5), np.log(2e6)))),
                                  nu=float(rng.uniform(0.40, 0.45)), density=1100.0, env_id=i))
        extra.append((rng, V, T, h))
        if progress is not None and (i + 1) % 256 == 0:
            progress(i + 1, n_env, "meshes")
    m.finalize()
    if progress is not None:
        progress(n_env, n_env, "finalized")
    m.gravity = (0.0, -9.81)
    tendons = []
    for hd, (rng, V, T, h) in zip(hs, extra):
        c = V[T].mean(1)
        ctr = np.array([0.0, V[:, 1].min() + R])
        taken = np.zeros(len(T), bool)
        for _ in range(int(rng.integers(1, 4))):
            for _try in range(50):
                p0 = ctr + rng.uniform(-0.5, 0.5, 2) * R
                rad = rng.uniform(0.12, 0.22) * R
                sel = np.linalg.norm(c - p0, axis=1) < rad
                near = np.linalg.norm(c - p0, axis=1) < rad + 3.0 * h
                if sel.sum() >= 4 and not (near & taken).any():
                    taken |= sel
                    m.add_affine_region(hd, sel)
                    break
        tl = []
        for _ in range(2):
            way = ctr + np.stack([np.linspace(-0.6, 0.6, 4), rng.uniform(-0.4, 0.4, 4)], 1) * R
            t = m.add_tendon(hd, through=way)
            m.set_tendon_tension(t, float(rng.uniform(0.0, 50.0)))
            tl.append(t)
        tendons.append(tl)
    if progress is not None:
        progress(n_env, n_env, "affine regions and tendons")
    return dict(model=m, dt=2e-3, command=lambda step: None, bodies=hs, tendons=tendons, size=2 * R,
                name="star_meshes")


def _table_beam(n):
    m = Model()
    m.damping = 50.0
    handles = [m.add_soft_body(np.ones((8, 80), np.uint8), physical_size=(0.040, 0.004), center=(0.0, 0.0),
                               E=1.0e6, nu=0.40, density=1100.0, env_id=i) for i in range(n)]
    m.finalize()
    cables = []
    for h in handles:
        X = m.rest_x(h)
        m.set_kinematic_constraint(h, np.flatnonzero(X[:, 0] < X[:, 0].min() + 1e-9))
        columns = np.unique(np.round(X[:, 0], 9))
        cables.append(m.add_tendon(h, through=[(cx, 0.001) for cx in columns]))
    tension = 10.0 + 20.0 * np.arange(n) / max(n - 1, 1)

    def command(step):
        for c, t in zip(cables, tension):
            m.set_tendon_tension(c, float(t))
    return m, dict(self_collision=False, friction=False), command, 1e-3, 5


def _table_strip(n):
    w = strip_on_ground(n)
    return w["model"], {}, w["command"], w["dt"], 30


def _table_grip(n):
    px = 0.5e-3
    m = Model()
    m.body_friction = 0.8
    m.add_ground(0.0, mu=0.5)
    pads = []
    for i in range(n):
        m.add_affine_body(np.ones((16, 16), np.uint8), physical_size=(16 * px, 16 * px),
                          center=(0.0, 8 * px + 0.9 * px / 8), density=600.0, env_id=i)
        pads.append([m.add_soft_body(np.ones((16, 8), np.uint8), physical_size=(8 * px, 16 * px),
                                     center=(side * 14 * px, 8 * px + 0.8e-3), E=1.0e6, nu=0.40, density=1100.0,
                                     env_id=i) for side in (-1, 1)])
    m.finalize()
    m.gravity = (0.0, -9.81)
    rest = []
    for pair in pads:
        for side, h in zip((-1, 1), pair):
            X = m.rest_x(h)
            back = X[:, 0].min() if side < 0 else X[:, 0].max()
            m.set_kinematic_constraint(h, np.flatnonzero(np.abs(X[:, 0] - back) < 1e-9))
            rest.append((side, h, X))
    close = 40

    def command(step):
        a, up = min(step / close, 1.0), 20.0e-6 * max(step - close, 0)
        for side, h, X in rest:
            Y = X.copy()
            Y[:, 0] -= side * 2.25 * px * a
            Y[:, 1] += up
            m.set_kinematic_target(h, Y)
    return m, {}, command, 1e-3, close + 5


TABLE_SCENES = {"beam": _table_beam, "strip": _table_strip, "grip": _table_grip}

CONFIGS = {"c1": strip_on_ground, "c2": random_blobs, "c3": pads_lift_block, "c4": pneumatic_fingers,
           "c5": star_meshes}
