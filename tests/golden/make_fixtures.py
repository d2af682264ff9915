"""Write the golden fixtures that pin the host scene layer's indexing.

Run once, in the container that holds the reference checkout (its warp-free host files only:
core/mesh.py and core/actuators.py, loaded by file path, plus tests/reference_static.py);
the outputs are committed as small .npz files next to this script and are what the tests
read. Nothing on the GPU machine runs this.

    python tests/golden/make_fixtures.py [/path/to/reference]
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

try:
    from .fixture_scenes import scenes, MESH_KEYS
except ImportError:                      # run as a script
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from fixture_scenes import scenes, MESH_KEYS

HERE = Path(__file__).resolve().parent


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod                      # dataclasses look their module up here
    spec.loader.exec_module(mod)
    return mod


def main(ref_root):
    ref = Path(ref_root)
    mesh = _load(ref / "core" / "mesh.py", "ref_core_mesh")
    act = _load(ref / "core" / "actuators.py", "ref_core_actuators")
    out = {}
    for name, bodies in scenes().items():
        me = mesh.build_scene([dict(b) for b in bodies])
        for k in MESH_KEYS:
            out[f"{name}/{k}"] = np.asarray(getattr(me, k))
    np.savez_compressed(HERE / "mesh_fixtures.npz", **out)

    # chamber edges of the ring's cavity and the exterior edges of the notch (pixel bodies)
    ring = scenes()["multi"][0]["mask_xy"].T
    lab, ids = act.enclosed_voids(ring)
    me = mesh.build_scene([dict(scenes()["multi"][0], env=0)])
    remap = me.grid_to_live_node_per_body[0]
    ny = ring.shape[0]

    def node_of(ix, iy):
        loc = int(remap[(ny + 1) * ix + iy])
        return -1 if loc < 0 else loc
    ea, eb = act.chamber_edges(ring, lab, ids[0], node_of)
    xa, xb = act.exterior_edges(ring, lab, ids, node_of)
    pts = np.array([[-3e-3, 2e-3], [-1e-3, 2.2e-3], [1.1e-3, 2.4e-3], [4e-3, 2.6e-3]])
    path = act.snap_path(pts, me.rest_pos, np.arange(me.rest_pos.shape[0]))
    np.savez_compressed(HERE / "actuator_fixtures.npz", labels=lab, ids=np.asarray(ids),
                        chamber_a=ea, chamber_b=eb, exterior_a=xa, exterior_b=xb,
                        tendon_points=pts, tendon_path=path)
    # the reference's independent float64 Newton oracle (tests/reference_static.py): elastic
    # energy and gradient on a randomly deformed blob, and the static equilibrium of a small
    # cantilever (left column fixed) under a dead load
    rs = _load(ref / "tests" / "reference_static.py", "ref_static")
    me = mesh.build_scene([dict(b) for b in scenes()["blob"]])
    rng = np.random.default_rng(11)
    x = me.rest_pos + 2e-4 * rng.standard_normal(me.rest_pos.shape)
    E, g, _ = rs.elastic(x, np.asarray(me.tri), np.asarray(me.D_m_inv), np.asarray(me.rest_area),
                         np.asarray(me.mu), np.asarray(me.lam), hess=False)
    cant = [dict(mask_xy=np.ones((4, 20), bool).T.copy(), size=(0.02, 0.004), center=(0.0, 0.0),
                 young=1e5, nu=0.4, rho=1000.0, env=0)]
    mc = mesh.build_scene(cant)
    fixed = np.nonzero(np.abs(mc.rest_pos[:, 0] - mc.rest_pos[:, 0].min()) < 1e-12)[0]
    fext = np.zeros_like(mc.rest_pos)
    fext[:, 1] = -np.asarray(mc.mass) * 9.81 * 20.0
    xs, res = rs.solve_static(mc, fixed, fext=fext, ramps=4, tol=1e-12)
    np.savez_compressed(HERE / "static_fixtures.npz", blob_x=x, blob_E=np.float64(E), blob_grad=g,
                        cant_fixed=fixed, cant_fext=fext, cant_x=xs, cant_residual=np.float64(res))
    print("wrote", HERE / "mesh_fixtures.npz", HERE / "actuator_fixtures.npz", HERE / "static_fixtures.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
