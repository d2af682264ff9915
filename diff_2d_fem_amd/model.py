"""Model: the scene container — bodies, colliders, contact parameters, kinematic targets,
actuators — with the reference's API (reference sim_model.py:46-1139).

    model = Model()
    model.add_ground(0.0, mu=0.8)
    h = model.add_soft_body(mask, physical_size=(0.054, 0.0045), center=(0, 0.003),
                            E=0.5e6, nu=0.40, density=1100.0)
    model.finalize()
    model.gravity = (0.0, -9.81)
    integ = ImplicitIntegrator()
    for t in range(T):
        integ.simulate(model, dt)
    x = model.get_body_x(h)

The Model is host bookkeeping only. Its device state lives in the solver an
ImplicitIntegrator attaches on the first simulate(); get_x and friends read it back.
"""
from __future__ import annotations

import hashlib
import warnings

import numpy as np

from . import geometry as G
from .config import SimConfig
from .mesh import build_scene, check_mesh, lame, meshes_overlap

# collider kinds (reference core/objects.py:22-24)
NONE, DISC, ELLIPSE, BOX, TRIANGLE, HALFSPACE = -1, 0, 1, 2, 3, 4
KIND_NAMES = {"disc": DISC, "ellipse": ELLIPSE, "box": BOX, "square": BOX, "triangle": TRIANGLE}


class Model:
    def __init__(self, device="cuda", config=None):
        if not str(device).startswith("cuda"):
            raise ValueError(f"this framework steps scenes on the GPU only (device='cuda'); "
                             f"got device={device!r}")
        self.cfg = SimConfig() if config is None else config
        self.device = device
        self._colliders = []
        self._bodies = []
        self._mesh_info = {}
        self._constrained = {}
        self.me = None
        self._solver = None
        self._targets = None
        self._node_range = []
        self._tri_range = []
        self._v0 = None
        self._x_host = None
        self._epoch = 0
        self.time = 0.0
        self._gravity = (0.0, 0.0)
        self._fext = None
        self._chambers = []
        self._tendons = []
        self._regions = []
        self._affine = []
        self._affine_of, self._regions_of = {}, {}
        self._springs = []
        self.body_contact = "ipc"
        self.body_friction = 0.0
        self.body_kappa = None
        self.rigid_young = 1.0e9
        self.rigid_nu = 0.30
        self._act_version = 0
        self._ctrl_version = 0

    # ---- contact / material attributes proxied onto the config
    def _p(name):
        return property(lambda s: getattr(s.cfg, name), lambda s, v: setattr(s.cfg, name, v))
    dhat = _p("ipc_d_hat")
    kappa = _p("obstacle_stiffness")
    self_stiffness = _p("contact_stiffness")
    mu = _p("obstacle_friction")
    epsv = _p("obstacle_friction_epsv")
    damping = _p("damping")
    del _p

    # ---- rigid colliders (reference sim_model.py:133-197)
    def add_disc(self, center, radius, mu=None):
        return self.add_object("disc", center, radius, mu=mu)

    def add_plane(self, normal, origin, mu=None):
        n = np.asarray(normal, np.float64)
        if not np.linalg.norm(n) > 0:
            raise ValueError("the plane normal must be non-zero")
        n = n / np.linalg.norm(n)
        return self._set_collider(HALFSPACE, origin, float(n[0]), float(n[1]), 0.0, mu)

    def add_ground(self, height=0.0, mu=None):
        return self.add_plane((0.0, 1.0), (0.0, height), mu=mu)

    def add_object(self, kind, center, rx, ry=None, corner_radius=3.0e-3, mu=None):
        k = KIND_NAMES[kind] if isinstance(kind, str) else int(kind)
        rx = float(rx)
        ry = rx if ry is None else float(ry)
        rref = float(corner_radius) if k in (BOX, TRIANGLE) else rx
        return self._set_collider(k, center, rx, ry, rref, mu)

    def _set_collider(self, kind, center, rx, ry, rref, mu):
        if self._solver is not None:
            raise RuntimeError("add colliders before the first simulate()")
        mu = float(self.cfg.obstacle_friction if mu is None else mu)
        self._colliders.append(dict(kind=int(kind), cx=float(center[0]), cy=float(center[1]),
                                    rx=rx, ry=ry, rref=rref, r=rx, mu=mu))
        return len(self._colliders) - 1

    def set_collider_center(self, index, center):
        c = self._colliders[index]
        c["cx"], c["cy"] = float(center[0]), float(center[1])
        if self._solver is not None:
            self._solver.set_collider(index, c)

    # ---- bodies (reference sim_model.py:199-338)
    def add_soft_body(self, mask, physical_size, center, E=None, nu=None, density=None,
                      env_id=0, mu=None):
        if self.me is not None:
            raise RuntimeError("add bodies before finalize()")
        mask = (np.asarray(mask) > 0.5).astype(np.uint8)
        if mask.ndim != 2 or not mask.any():
            raise ValueError("the mask must be a 2-D array with at least one solid pixel")
        if E is not None and not np.isscalar(E):
            E = np.asarray(E, np.float64)
            if E.shape != mask.shape:
                raise ValueError(f"stiffness map is {E.shape}, the mask is {mask.shape}")
            if (E[mask > 0] <= 0).any():
                raise ValueError("stiffness map must be positive wherever the mask is solid")
        if int(env_id) < 0:
            raise ValueError("env_id must be >= 0")
        c = self.cfg
        self._bodies.append(dict(
            mask=mask, size=tuple(float(v) for v in physical_size),
            center=tuple(float(v) for v in center),
            E=(c.young if E is None else (float(E) if np.isscalar(E) else E)),
            nu=float(c.nu if nu is None else nu),
            rho=float(c.rho_mass if density is None else density),
            env=int(env_id), mu=self._check_mu(mu)))
        return len(self._bodies) - 1

    def add_affine_body(self, mask, physical_size, center, density=None, env_id=0, E=None,
                        nu=None, mu=None):
        h = self.add_soft_body(mask, physical_size, center,
                               E=(self.rigid_young if E is None else E),
                               nu=(self.rigid_nu if nu is None else nu),
                               density=density, env_id=env_id, mu=mu)
        if not np.isscalar(self._bodies[h]["E"]):
            raise ValueError("a rigid body takes a scalar stiffness")
        self._bodies[h]["affine"] = True
        return h

    def add_kinematic_body(self, mask, physical_size, center, env_id=0, mu=None):
        h = self.add_soft_body(mask, physical_size, center, env_id=env_id, mu=mu)
        self._bodies[h]["kinematic"] = True
        return h

    def add_soft_mesh(self, vertices, triangles, E=None, nu=None, density=None, env_id=0, mu=None):
        if self.me is not None:
            raise RuntimeError("add bodies before finalize()")
        info = self._check_mesh(vertices, triangles)
        m = len(info["triangles"])
        if E is not None and not np.isscalar(E):
            E = np.asarray(E, np.float64)
            if E.shape != (m,):
                raise ValueError(f"stiffness per triangle is {E.shape}; the mesh has {m} triangles")
            if not (np.isfinite(E).all() and (E > 0).all()):
                raise ValueError("stiffness must be positive and finite on every triangle")
        if int(env_id) < 0:
            raise ValueError("env_id must be >= 0")
        c = self.cfg
        V = info["vertices"]
        self._bodies.append(dict(
            mesh=True, vertices=V, triangles=info["triangles"], h=info["h"],
            h_cfl=info["h_cfl"], L_max=info["L_max"],
            center=tuple(float(v) for v in 0.5 * (V.min(0) + V.max(0))),
            E=(c.young if E is None else (float(E) if np.isscalar(E) else E)),
            nu=float(c.nu if nu is None else nu),
            rho=float(c.rho_mass if density is None else density),
            env=int(env_id), mu=self._check_mu(mu)))
        return len(self._bodies) - 1

    def add_affine_mesh(self, vertices, triangles, density=None, env_id=0, E=None, nu=None, mu=None):
        h = self.add_soft_mesh(vertices, triangles, E=(self.rigid_young if E is None else E),
                               nu=(self.rigid_nu if nu is None else nu), density=density,
                               env_id=env_id, mu=mu)
        if not np.isscalar(self._bodies[h]["E"]):
            raise ValueError("a rigid body takes a scalar stiffness")
        self._bodies[h]["affine"] = True
        return h

    def add_kinematic_mesh(self, vertices, triangles, env_id=0, mu=None):
        h = self.add_soft_mesh(vertices, triangles, env_id=env_id, mu=mu)
        self._bodies[h]["kinematic"] = True
        return h

    @staticmethod
    def _check_mu(mu):
        if mu is None:
            return None
        mu = float(mu)
        if not (np.isfinite(mu) and mu >= 0.0):
            raise ValueError(f"mu must be a finite number >= 0, not {mu}")
        return mu

    def _body_mu_per_node(self):
        if all(b.get("mu") is None for b in self._bodies):
            return None
        out = np.zeros(self.n_nodes, np.float64)
        for h, b in enumerate(self._bodies):
            out[self._node_range[h]] = self.body_friction if b.get("mu") is None else b["mu"]
        return out

    def _check_mesh(self, vertices, triangles):
        V = np.ascontiguousarray(vertices, np.float64)
        T = np.ascontiguousarray(triangles)
        key = (V.shape, T.shape, str(T.dtype), hashlib.sha1(V.tobytes() + T.tobytes()).hexdigest())
        if key not in self._mesh_info:
            info = check_mesh(V, T)
            s = np.sqrt(2.0 * info["area"])
            if s.max() > 10.0 * s.min():
                warnings.warn(f"this mesh's elements range from {s.min() * 1e3:.3g} to "
                              f"{s.max() * 1e3:.3g} mm; the iterations a step needs follow the "
                              f"smallest element.", stacklevel=4)
            self._mesh_info[key] = info
        return self._mesh_info[key]

    def _is_mesh(self, handle):
        return bool(self._bodies[handle].get("mesh"))

    def set_kinematic_pose(self, handle, translation=(0.0, 0.0), angle=0.0, about=None):
        self.set_kinematic_target(handle, self.rigid_motion(handle, translation, angle, about))

    # ---- build (reference sim_model.py:391-429)
    def finalize(self):
        if not self._bodies:
            raise ValueError("add at least one soft body before finalize()")
        B = self._bodies
        dx = self.dx
        r_excl = max(2.0 * dx, 16.0 * (self.cfg.ipc_d_hat if self.cfg.ipc_d_hat is not None else dx / 8.0))

        def geom(b):
            if b.get("mesh"):
                return dict(vertices=b["vertices"], triangles=b["triangles"], excl_radius=r_excl,
                            checked=True)
            return dict(mask_xy=b["mask"].T.copy(), size=b["size"], center=b["center"])
        self.me = build_scene([dict(**geom(b), young=(b["E"] if np.isscalar(b["E"]) else 1.0),
                                    nu=b["nu"], rho=b["rho"], env=b["env"]) for b in B])
        me = self.me
        self._node_range = [np.arange(*me.body_node_range[h]) for h in range(len(B))]
        self._tri_range = [np.arange(*me.body_tri_range[h]) for h in range(len(B))]
        for h, b in enumerate(B):
            if np.isscalar(b["E"]):
                continue
            tris = self._tri_range[h]
            scale = (b["E"] if b.get("mesh") else
                     G.pixel_field_on_elements(b["E"], me.tri_quad, tris, b["mask"].shape[0]))
            me.mu[tris] = me.mu[tris] * scale
            me.lam[tris] = me.lam[tris] * scale
        self._check_layout()
        self._scale_contact()
        self._targets = me.rest_pos.astype(np.float64).copy()
        self._x_host = me.rest_pos.astype(np.float64).copy()
        self._v0 = np.zeros_like(self._x_host)
        for h, b in enumerate(B):
            if b.get("affine"):
                self._add_affine(h, self._tri_range[h])
            if b.get("kinematic"):
                self._constrained[h] = self._node_range[h].copy()
        return self

    def add_affine_region(self, handle, region, E=None, nu=None):
        self._need_mesh()
        if self._solver is not None:
            raise RuntimeError("add rigid regions before the first simulate()")
        elems = self._region_elements(handle, region)
        mu, lam = lame(self.rigid_young if E is None else float(E),
                       self.rigid_nu if nu is None else float(nu))
        self.me.mu[elems] = mu
        self.me.lam[elems] = lam
        return self._add_affine(handle, elems)

    def _add_affine(self, handle, elems):
        """Register a rigid region; its boundary nodes stop contacting one another
        (reference sim_model.py:454-471)."""
        me = self.me
        mine = self._affine_of.setdefault(handle, [])
        if mine and np.intersect1d(elems, np.concatenate(mine)).size:
            raise ValueError("rigid regions must not overlap")
        mine.append(np.asarray(elems, np.int64))
        nodes = np.unique(np.asarray(me.tri)[elems].reshape(-1))
        env = self._bodies[handle]["env"]
        bn = np.asarray(me.boundary_nodes)
        lo, hi = int(me.boundary_starts[env]), int(me.boundary_starts[env + 1])
        loc = np.nonzero(np.isin(bn[lo:hi], nodes))[0]
        ex = me.excluded_pairs_per_env[env]
        ex[np.ix_(loc, loc)] = True
        o = int(me.excl_offset[env])
        me.excl_flat[o:o + ex.size] = ex.reshape(-1).astype(np.uint8)
        self._affine.append(dict(body=handle, elems=np.asarray(elems, np.int64), nodes=nodes))
        return len(self._affine) - 1

    def affine_state(self, index):
        s = self._solver
        if s is None or s.affine is None:
            a = self._affine[index]
            m = np.asarray(self.me.mass)[a["nodes"]]
            c = (m[:, None] * self.me.rest_pos[a["nodes"]]).sum(0) / m.sum()
            return dict(position=c, angle=0.0, A=np.eye(2))
        t, A = s.affine_state()
        return dict(position=t[index], angle=float(np.arctan2(A[index][1, 0] - A[index][0, 1],
                                                              A[index][0, 0] + A[index][1, 1])),
                    A=A[index])

    def _scale_contact(self):
        """Contact parameters left unset, from the scene's scales (reference
        sim_model.py:487-497): E the softest soft body, dx the largest element size."""
        c = self.cfg
        soft = [b for b in self._bodies if not b.get("affine") and not b.get("kinematic")]
        E = min((float(b["E"]) if np.isscalar(b["E"]) else
                 float(b["E"].min()) if b.get("mesh") else float(b["E"][b["mask"] > 0].min())
                 for b in soft), default=1.0e6)
        if c.ipc_d_hat is None:
            c.ipc_d_hat = self.dx / 8.0
        if c.obstacle_stiffness is None:
            c.obstacle_stiffness = 4.0e-3 * E * self.dx / c.ipc_d_hat
        if c.contact_stiffness is None:
            c.contact_stiffness = 0.2 * E
        if self.body_kappa is None:
            self.body_kappa = 0.1 * E * self.dx / c.ipc_d_hat

    def _check_layout(self):
        """Bodies of one scene must start apart (reference sim_model.py:499-536)."""
        envs = {}
        for h, b in enumerate(self._bodies):
            envs.setdefault(b["env"], []).append(h)
        memo = {}
        for env, hs in envs.items():
            for i, a in enumerate(hs):
                for c in hs[i + 1:]:
                    if self._is_mesh(a) or self._is_mesh(c):
                        key = (self._geom_key(a), self._geom_key(c))
                        if key not in memo:
                            memo[key] = self._overlap_mesh(a, c)
                        hit = memo[key]
                    else:
                        key = (self._geom_key(a), self._geom_key(c))
                        if key not in memo:
                            memo[key] = self._overlap(a, c) or self._overlap(c, a)
                        hit = memo[key]
                    if hit:
                        raise ValueError(
                            f"bodies {a} and {c} overlap in environment {env}. Bodies of one "
                            f"environment collide with each other and must start apart; give "
                            f"the members of a batch their own environment with env_id=i.")
            if any(self._is_mesh(h) for h in hs):
                sz = [self.element_size(h) for h in hs]
                if max(sz) > 1.25 * min(sz):
                    warnings.warn(f"environment {env} mixes element sizes from {min(sz) * 1e3:.3g} "
                                  f"to {max(sz) * 1e3:.3g} mm", stacklevel=3)
                continue
            px = [self.pixel_size(h)[0] for h in hs]
            if max(px) > 1.25 * min(px):
                warnings.warn(f"environment {env} mixes pixel sizes from {min(px) * 1e3:.3g} to "
                              f"{max(px) * 1e3:.3g} mm", stacklevel=3)

    def _geom_key(self, h):
        b = self._bodies[h]
        if b.get("mesh"):
            return ("mesh", id(b["vertices"]), id(b["triangles"]))
        return ("mask", b["mask"].shape, b["mask"].tobytes(), b["size"], b["center"])

    def _body_geometry(self, h):
        me, nodes = self.me, self._node_range[h]
        s, e = int(nodes[0]), int(nodes[-1]) + 1
        return np.asarray(me.rest_pos[s:e], np.float64), np.asarray(me.tri)[self._tri_range[h]] - s

    def _overlap_mesh(self, a, c):
        return meshes_overlap(*self._body_geometry(a), *self._body_geometry(c))

    def _overlap(self, a, c):
        """Does the centre of any solid pixel of body a fall in a solid pixel of body c?"""
        A, C = self._bodies[a], self._bodies[c]
        (aw, ah), (ax, ay) = A["size"], A["center"]
        (cw, ch), (cx, cy) = C["size"], C["center"]
        if abs(ax - cx) >= 0.5 * (aw + cw) or abs(ay - cy) >= 0.5 * (ah + ch):
            return False
        H, W = A["mask"].shape
        row, col = np.nonzero(A["mask"])
        px = ax - 0.5 * aw + (col + 0.5) * aw / W
        py = ay - 0.5 * ah + (row + 0.5) * ah / H
        Hc, Wc = C["mask"].shape
        jc = np.floor((px - (cx - 0.5 * cw)) / (cw / Wc)).astype(int)
        jr = np.floor((py - (cy - 0.5 * ch)) / (ch / Hc)).astype(int)
        ok = (jc >= 0) & (jc < Wc) & (jr >= 0) & (jr < Hc)
        return bool(C["mask"][jr[ok], jc[ok]].any())

    # ---- layout queries
    @property
    def n_bodies(self):
        return len(self._bodies)

    @property
    def n_envs(self):
        return int(self.me.n_envs) if self.me is not None else \
            (max(b["env"] for b in self._bodies) + 1 if self._bodies else 0)

    @property
    def n_nodes(self):
        return int(self.me.rest_pos.shape[0])

    @property
    def n_elements(self):
        return int(self.me.tri.shape[0])

    def body_env(self, handle):
        return self._bodies[handle]["env"]

    def body_nodes(self, handle):
        return self._node_range[handle].copy()

    def body_elements(self, handle):
        return self._tri_range[handle].copy()

    def pixel_size(self, handle):
        if self._is_mesh(handle):
            raise ValueError(f"body {handle} is a mesh body: it has no pixels")
        b = self._bodies[handle]
        H, W = b["mask"].shape
        return b["size"][0] / W, b["size"][1] / H

    def element_size(self, handle):
        b = self._bodies[handle]
        return float(b["h"]) if b.get("mesh") else self.pixel_size(handle)[0]

    def _cfl_length(self, handle):
        b = self._bodies[handle]
        return float(b["h_cfl"]) if b.get("mesh") else min(self.pixel_size(handle))

    def element_areas(self, handle):
        return np.asarray(self.me.rest_area, np.float64)[self._tri_range[handle]].copy()

    def rest_x(self, handle):
        return self.me.rest_pos[self._node_range[handle]].copy()

    def to_elements(self, handle, field):
        b = self._bodies[handle]
        field = np.asarray(field, np.float64)
        if b.get("mesh"):
            m = len(self._tri_range[handle])
            if field.shape != (m,):
                raise ValueError(f"field is {field.shape}; mesh body {handle} has {m} triangles")
            return field.copy()
        if field.shape != b["mask"].shape:
            raise ValueError(f"field is {field.shape}, the mask is {b['mask'].shape}")
        return G.pixel_field_on_elements(field, self.me.tri_quad, self._tri_range[handle],
                                         b["mask"].shape[0])

    def to_pixels(self, handle, values):
        if self._is_mesh(handle):
            raise ValueError(f"body {handle} is a mesh body: to_pixels needs a pixel body")
        H, W = self._bodies[handle]["mask"].shape
        q = np.asarray(self.me.tri_quad)[self._tri_range[handle]]
        out = np.zeros((H, W))
        np.add.at(out, (q % H, q // H), np.asarray(values, np.float64))
        return out

    def rigid_motion(self, handle, translation=(0.0, 0.0), angle=0.0, about=None):
        x = self.rest_x(handle)
        c = np.asarray(self._bodies[handle]["center"] if about is None else about, np.float64)
        ca, sa = np.cos(angle), np.sin(angle)
        R = np.array([[ca, -sa], [sa, ca]])
        return (x - c) @ R.T + c + np.asarray(translation, np.float64)

    # ---- kinematic boundary conditions (reference sim_model.py:686-703)
    def set_kinematic_constraint(self, handle, node_indices):
        self._constrained[handle] = self._node_range[handle][np.asarray(node_indices, int)]

    def set_kinematic_target(self, handle, targets):
        self._targets[self._node_range[handle]] = np.asarray(targets, np.float64)

    def get_body_x(self, handle):
        return self.get_x()[self._node_range[handle]]

    def get_x(self):
        if self._solver is not None:
            self._x_host = self._solver.get_x()
        return self._x_host.copy()

    def get_body_v(self, handle):
        v = self._v0 if self._solver is None else self._solver.get_v()
        return np.asarray(v)[self._node_range[handle]].copy()

    def set_body_velocity(self, handle, velocity=(0.0, 0.0), omega=0.0, about=None):
        nodes = self._node_range[handle]
        x = self.get_x()[nodes].astype(np.float64)
        if about is not None:
            c = np.asarray(about, np.float64)
        elif self._is_mesh(handle):
            m = np.asarray(self.me.mass, np.float64)[nodes]
            c = (m[:, None] * x).sum(0) / m.sum()
        else:
            c = x.mean(0)
        r = x - c
        v = np.asarray(velocity, np.float64)[None, :] + omega * np.stack([-r[:, 1], r[:, 0]], 1)
        self._v0[nodes] = v
        if self._solver is not None:
            self._solver.set_v_rows(nodes, v)

    def set_body_x(self, handle, x, v=None):
        nodes = self._node_range[handle]
        x = np.asarray(x, np.float64).reshape(len(nodes), 2)
        self._x_host[nodes] = x
        self._targets[nodes] = x
        if v is not None:
            self._v0[nodes] = np.asarray(v, np.float64).reshape(len(nodes), 2)
        s = self._solver
        if s is not None:
            s.set_nodes(int(nodes[0]), x, None if v is None else self._v0[nodes])
            s.sync_affine([i for i, a in enumerate(self._affine) if a["body"] == handle])

    def set_body_pose(self, handle, translation=(0.0, 0.0), angle=0.0, about=None,
                      velocity=(0.0, 0.0), omega=0.0):
        self.set_body_x(handle, self.rigid_motion(handle, translation, angle, about))
        self.set_body_velocity(handle, velocity, omega)

    def reset(self, env_ids=None):
        """Back to rest at zero velocity (reference sim_model.py:761-788)."""
        rest = self.me.rest_pos.astype(np.float64)
        s = self._solver
        if env_ids is None:
            self._targets = rest.copy()
            self._x_host = rest.copy()
            self._v0 = np.zeros_like(rest)
            self._epoch += 1
            self.time = 0.0
            if s is not None:
                s.reset_all(rest)
            return
        pick = np.zeros(self.n_envs, np.int32)
        pick[np.atleast_1d(env_ids)] = 1
        nodes = pick[np.asarray(self.me.node2env)] != 0
        self._targets[nodes] = rest[nodes]
        self._x_host[nodes] = rest[nodes]
        self._v0[nodes] = 0.0
        self._epoch += 1
        if s is not None:
            s.reset_envs(pick)
            s.sync_affine([i for i, a in enumerate(self._affine) if pick[self._bodies[a["body"]]["env"]]])

    # ---- actuation (reference sim_model.py:790-1096)
    def _need_mesh(self):
        if self.me is None:
            raise RuntimeError("call finalize() before adding actuators")

    def _node_of(self, handle):
        return lambda ix, iy: self.me.body_node_index(handle, int(ix), int(iy))

    @property
    def gravity(self):
        return self._gravity

    @gravity.setter
    def gravity(self, g):
        self._gravity = (float(g[0]), float(g[1]))
        self._ctrl_version += 1

    def set_external_force(self, handle, force):
        self._need_mesh()
        if self._fext is None:
            self._fext = np.zeros((self.me.rest_pos.shape[0], 2), np.float64)
        self._fext[self._node_range[handle]] = np.asarray(force, np.float64)
        self._ctrl_version += 1

    def add_pressure_chambers(self, handle, min_pixels=1):
        self._need_mesh()
        if self._is_mesh(handle):
            holes = [lp for lp in self._mesh_loops(handle) if lp[2] > 0]
            rest = self.me.rest_pos

            def low(lp):
                return tuple(rest[lp[0][np.argmin(rest[lp[0], 1])]][::-1])
            return [self._register_chamber(handle, ea, eb, A, sign=1.0)
                    for ea, eb, A in sorted(holes, key=low)]
        mask = self._bodies[handle]["mask"]
        lab, ids = G.enclosed_voids(mask)
        ids = [k for k in ids if (lab == k).sum() >= min_pixels]
        ids.sort(key=lambda k: tuple(np.argwhere(lab == k)[0]))
        return [self._add_chamber(handle, lab, k) for k in ids]

    def add_pressure_chamber(self, handle, at):
        self._need_mesh()
        if self._is_mesh(handle):
            rest = np.asarray(self.me.rest_pos, np.float64)
            for ea, eb, A in self._mesh_loops(handle):
                if A > 0 and G.point_in_loop(at, rest, ea, eb):
                    return self._register_chamber(handle, ea, eb, A, sign=1.0)
            raise ValueError(f"point {tuple(at)} is not inside a hole of this mesh body")
        b = self._bodies[handle]
        mask = b["mask"]
        px = G.pixel_of(at, b["size"], b["center"], mask.shape)
        if px is None or mask[px] > 0:
            raise ValueError(f"point {tuple(at)} is not inside a void of this body")
        lab, ids = G.enclosed_voids(mask)
        if lab[px] not in ids:
            raise ValueError(f"the void at {tuple(at)} is open to the outside, not a sealed cavity")
        return self._add_chamber(handle, lab, int(lab[px]))

    def add_ambient_pressure(self, handle):
        self._need_mesh()
        if self._is_mesh(handle):
            out = [lp for lp in self._mesh_loops(handle) if lp[2] < 0]
            return self._register_chamber(handle, np.concatenate([lp[0] for lp in out]),
                                          np.concatenate([lp[1] for lp in out]),
                                          -sum(lp[2] for lp in out), sign=-1.0)
        mask = self._bodies[handle]["mask"]
        lab, ids = G.enclosed_voids(mask)
        ea, eb = G.exterior_edges(mask, lab, ids, self._node_of(handle))
        n_px = int((mask > 0).sum() + sum((lab == k).sum() for k in ids))
        pw, ph = self.pixel_size(handle)
        return self._register_chamber(handle, ea, eb, n_px * pw * ph, sign=-1.0)

    def _add_chamber(self, handle, lab, label):
        ea, eb = G.chamber_edges(self._bodies[handle]["mask"], lab, label, self._node_of(handle))
        pw, ph = self.pixel_size(handle)
        return self._register_chamber(handle, ea, eb, int((lab == label).sum()) * pw * ph, sign=1.0)

    def _mesh_loops(self, handle):
        nodes = self._node_range[handle]
        s0 = int(nodes[0])
        tri = np.asarray(self.me.tri)[self._tri_range[handle]] - s0
        rest = np.asarray(self.me.rest_pos, np.float64)[nodes]
        loops = [(ea + s0, eb + s0, A) for ea, eb, A in G.mesh_boundary_loops(tri, rest)]
        outline = -sum(A for _, _, A in loops if A < 0)
        inside = self.element_areas(handle).sum() + sum(A for _, _, A in loops if A > 0)
        if abs(outline - inside) > 1e-9 * outline:
            raise RuntimeError(f"mesh body {handle}: its outline encloses {outline:.9e} m^2, its "
                               f"material and holes {inside:.9e}")
        return loops

    def _register_chamber(self, handle, ea, eb, expect, sign):
        rest = self.me.rest_pos
        ref = rest[np.concatenate([ea, eb])].mean(0)
        got = sign * G.loop_area(rest, ea, eb, ref)
        if not expect > 0 or abs(got - expect) > 1e-6 * expect:
            raise RuntimeError(f"pressure boundary encloses {got:.6e} m^2, expected {expect:.6e}")
        self._chambers.append(dict(body=handle, ea=ea, eb=eb, ref=ref, area0=float(expect),
                                   sign=sign, pressure=0.0))
        self._act_version += 1
        return len(self._chambers) - 1

    def set_pressure(self, chamber, pressure):
        self._chambers[chamber]["pressure"] = float(pressure)
        self._ctrl_version += 1

    def chamber_area(self, chamber):
        c = self._chambers[chamber]
        return c["sign"] * G.loop_area(self.get_x(), c["ea"], c["eb"], c["ref"])

    def add_tendon(self, handle, through=None, nodes=None, stiffness=None):
        self._need_mesh()
        ids = self._node_range[handle]
        if (through is None) == (nodes is None):
            raise ValueError("give exactly one of `through` (points) or `nodes` (indices)")
        path = ids[np.asarray(nodes, int)] if nodes is not None else \
            G.snap_path(through, self.me.rest_pos[ids], ids)
        if len(path) < 2:
            raise ValueError("a tendon needs at least two distinct eyelets")
        if stiffness is not None and stiffness <= 0:
            raise ValueError("stiffness must be positive (or None for a tension-controlled cable)")
        self._tendons.append(dict(body=handle, nodes=np.asarray(path, np.int64),
                                  stiffness=0.0 if stiffness is None else float(stiffness),
                                  tension=0.0, length=G.path_length(self.me.rest_pos, path)))
        self._act_version += 1
        return len(self._tendons) - 1

    def set_tendon_tension(self, tendon, tension):
        t = self._tendons[tendon]
        if t["stiffness"] > 0:
            raise ValueError("this tendon is length-controlled; use set_tendon_length")
        t["tension"] = float(tension)
        self._ctrl_version += 1

    def set_tendon_length(self, tendon, length):
        t = self._tendons[tendon]
        if t["stiffness"] <= 0:
            raise ValueError("this tendon is tension-controlled; use set_tendon_tension")
        t["length"] = float(length)
        self._ctrl_version += 1

    def tendon_length(self, tendon):
        return G.path_length(self.get_x(), self._tendons[tendon]["nodes"])

    def tendon_tension(self, tendon):
        t = self._tendons[tendon]
        if t["stiffness"] > 0:
            return t["stiffness"] * max(self.tendon_length(tendon) - t["length"], 0.0)
        return max(t["tension"], 0.0)

    def add_spring(self, handle_a, point_a, handle_b=None, point_b=None, stiffness=None,
                   rest_length=None):
        self._need_mesh()
        ia = self._nearest(handle_a, point_a)
        if handle_b is None:
            ib = -1
            anchor = np.asarray(self.me.rest_pos[ia] if point_b is None else point_b, np.float64)
        else:
            ib = self._nearest(handle_b, point_a if point_b is None else point_b)
            if self.body_env(handle_a) != self.body_env(handle_b):
                raise ValueError("a spring cannot join bodies of different environments")
            if ib == ia:
                raise ValueError("both ends of the spring snapped to the same node")
            anchor = np.asarray(self.me.rest_pos[ib], np.float64)
        L = float(np.linalg.norm(self.me.rest_pos[ia] - anchor)) if rest_length is None \
            else float(rest_length)
        if handle_b is not None and L <= 0.0:
            raise ValueError("a spring between two bodies needs a rest length")
        k = float(self.rigid_young if stiffness is None else stiffness)
        if k <= 0 or L < 0:
            raise ValueError("spring stiffness must be positive and its rest length non-negative")
        self._springs.append(dict(a=int(ia), b=int(ib), stiffness=k, length=L, anchor=anchor,
                                  env=self.body_env(handle_a)))
        self._act_version += 1
        return len(self._springs) - 1

    def add_pin(self, handle, point, handle_b=None, stiffness=None):
        if handle_b is not None:
            raise NotImplementedError("a pin between two bodies is not supported: pin to the "
                                      "world, or use add_spring with a rest length")
        return self.add_spring(handle, point, None, point, stiffness, rest_length=0.0)

    def set_spring_length(self, spring, length):
        self._springs[spring]["length"] = float(length)
        self._ctrl_version += 1

    def set_spring_anchor(self, spring, point):
        if self._springs[spring]["b"] >= 0:
            raise ValueError("this spring joins two bodies and has no world anchor")
        self._springs[spring]["anchor"] = np.asarray(point, np.float64)
        self._ctrl_version += 1

    def spring_force(self, spring):
        sp_ = self._springs[spring]
        x = self.get_x().astype(np.float64)
        ln = float(np.linalg.norm(x[sp_["a"]] - (x[sp_["b"]] if sp_["b"] >= 0 else sp_["anchor"])))
        return sp_["stiffness"] * (ln - sp_["length"]), ln

    def _nearest(self, handle, point):
        ids = self._node_range[handle]
        d = np.linalg.norm(self.me.rest_pos[ids] - np.asarray(point, np.float64)[None, :], axis=1)
        return int(ids[d.argmin()])

    def add_active_region(self, handle, region, fiber=None):
        self._need_mesh()
        elems = self._region_elements(handle, region)
        mine = self._regions_of.setdefault(handle, [])
        if mine and np.intersect1d(elems, np.concatenate(mine)).size:
            raise ValueError("active regions must not overlap")
        mine.append(np.asarray(elems, np.int64))
        self._regions.append(dict(body=handle, elems=elems, fiber=fiber, A=np.eye(2)))
        self._act_version += 1
        return len(self._regions) - 1

    def set_activation(self, region, stretch, transverse=1.0):
        r = self._regions[region]
        if stretch <= 0 or transverse <= 0:
            raise ValueError("stretch ratios must be positive")
        r["A"] = G.activation_matrix(stretch, transverse, r["fiber"])
        self._ctrl_version += 1

    def all_constrained(self):
        if not self._constrained:
            return np.zeros(0, np.int64)
        return np.concatenate([self._constrained[h] for h in sorted(self._constrained)])

    def _region_elements(self, handle, region):
        b = self._bodies[handle]
        if not b.get("mesh"):
            region = np.asarray(region) > 0
            shape = b["mask"].shape
            if region.shape != shape:
                raise ValueError(f"region is {region.shape}, the mask is {shape}")
            elems = G.region_elements(region, self.me.tri_quad, self._tri_range[handle], shape[0])
            if len(elems) == 0:
                raise ValueError("the region contains no solid pixel of this body")
            return elems
        tris = self._tri_range[handle]
        m = len(tris)
        r = np.asarray(region)
        if r.dtype == bool:
            if r.shape != (m,):
                raise ValueError(f"region is a boolean array of shape {r.shape}; mesh body "
                                 f"{handle} has {m} triangles")
            loc = np.nonzero(r)[0]
        else:
            if r.ndim != 1 or not np.issubdtype(r.dtype, np.integer):
                raise ValueError(f"region of a mesh body: ({m},) booleans or triangle indices")
            loc = np.unique(r)
            if len(loc) != len(r) or (len(loc) and (loc[0] < 0 or loc[-1] >= m)):
                raise ValueError(f"triangle indices must be distinct and in 0..{m - 1}")
        if len(loc) == 0:
            raise ValueError("the region contains no triangle of this body")
        return tris[loc]

    @property
    def dx(self):
        """Largest element size over the bodies: sets the thickness of contact."""
        return max(self.element_size(h) for h in range(len(self._bodies)))
