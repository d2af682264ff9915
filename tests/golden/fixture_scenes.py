"""The scenes whose host-side meshes are pinned by the golden fixtures (mesh_fixtures.npz,
actuator_fixtures.npz): name -> list of body dicts for build_scene. Shared by the fixture
generator and the tests that read the fixtures."""
from __future__ import annotations

import numpy as np


def scenes():
    """The scenes whose meshes are pinned: name -> list of body dicts for build_scene."""
    rng = np.random.default_rng(7)
    blob = rng.random((9, 14)) > 0.35
    blob[4, :] = True
    ring = np.ones((7, 9), bool)
    ring[2:5, 3:6] = False                      # a sealed cavity
    notch = np.ones((6, 10), bool)
    notch[3:, 4:6] = False                      # a slot open to the top
    # a small jittered mesh body
    xs, ys = np.meshgrid(np.linspace(0, 4e-3, 5), np.linspace(0, 2e-3, 3), indexing="ij")
    V = np.stack([xs.ravel(), ys.ravel()], 1)
    V[6] += [1.3e-4, -0.7e-4]
    T = []
    for i in range(4):
        for j in range(2):
            a = i * 3 + j
            T += [[a, a + 1, a + 4], [a, a + 4, a + 3]]
    T = np.asarray(T)
    return {
        # the strip of the quickstart: mask ones(6, 72), (H, W) with row 0 at the bottom
        "strip": [dict(mask_xy=np.ones((6, 72), bool).T.copy(), size=(0.054, 0.0045),
                       center=(0.0, 0.0030), young=0.5e6, nu=0.40, rho=1100.0, env=0)],
        "blob": [dict(mask_xy=blob.T.copy(), size=(14e-3, 9e-3), center=(0.01, 0.02),
                      young=2e5, nu=0.45, rho=1000.0, env=0)],
        # three scenes, two bodies in scene 1 given out of scene order
        "multi": [dict(mask_xy=ring.T.copy(), size=(9e-3, 7e-3), center=(0.0, 0.005),
                       young=3e5, nu=0.45, rho=1100.0, env=1),
                  dict(mask_xy=notch.T.copy(), size=(10e-3, 6e-3), center=(0.0, 0.004),
                       young=1e6, nu=0.40, rho=900.0, env=0),
                  dict(mask_xy=np.ones((4, 4), bool), size=(4e-3, 4e-3), center=(0.012, 0.003),
                       young=1e9, nu=0.30, rho=600.0, env=1),
                  dict(mask_xy=notch.T.copy(), size=(10e-3, 6e-3), center=(0.0, 0.004),
                       young=1e6, nu=0.40, rho=900.0, env=2)],
        "meshbody": [dict(vertices=V, triangles=T, excl_radius=2.5e-3, young=1e6, nu=0.4,
                          rho=1000.0, env=0)],
    }


MESH_KEYS = ("rest_pos", "tri", "D_m_inv", "rest_area", "mu", "lam", "mass", "node2env",
             "tri2env", "boundary_nodes", "boundary2env", "boundary_edges", "node_starts",
             "tri_starts", "boundary_starts", "boundary_per_env", "excl_flat", "excl_offset",
             "tri_quad", "body_node_range", "body_tri_range")
