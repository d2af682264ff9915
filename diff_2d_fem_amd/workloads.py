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
                    mu 0.6, ground mu 0.6, gravity, dt 1 ms. The jaws close at 1 mm per step to
                    a quarter pixel past touching, then lift at 20 um per step.
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
    rest = []
    for pair in pads:
        r = []
        for s, p in zip((-1, 1), pair):
            X = m.rest_x(p)
            back = X[:, 0].min() if s < 0 else X[:, 0].max()
            m.set_kinematic_constraint(p, np.where(np.abs(X[:, 0] - back) < 1e-9)[0])
            r.append(X)
        rest.append(r)
    close = np.array([2.25 * px for px in pxs])             # to a quarter pixel past touching

    lift = np.array([8.0 * px for px in pxs])

    def command(step):
        for i, (pair, xs) in enumerate(zip(pads, rest)):
            c = close[i] * min(step / n_close, 1.0)
            up = lift[i] * min(max(step - n_close, 0) / n_lift, 1.0)
            for s, p, X in zip((-1, 1), pair, xs):
                Y = X.copy()
                Y[:, 0] -= s * c
                Y[:, 1] += up
                m.set_kinematic_target(p, Y)
    return dict(model=m, dt=1e-3, command=command, bodies=blocks, pads=pads,
                size=float(np.mean([16 * px for px in pxs])),
                sides=np.array([16 * px for px in pxs]), name="pads_lift_block")


CONFIGS = {"c1": strip_on_ground, "c2": random_blobs, "c3": pads_lift_block}
