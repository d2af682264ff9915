"""Model: the scene container — bodies, colliders, contact parameters, kinematic targets and
actuators — with the reference's API (reference sim_model.py:46-1139: the same methods,
arguments, defaults and exceptions).

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

The Model keeps host records only (scene_spec): BodySpec per body, ColliderSpec per collider,
RigidRegionSpec per rigid piece, and an Actuation with the chambers, tendons, springs, active
regions, gravity and applied forces. finalize() lays the bodies out as one flat mesh
(mesh.build_scene). The state lives in the device solver an ImplicitIntegrator attaches on the
first simulate(); get_x and friends read it back.

Environments: every body belongs to one (env_id, default 0). Bodies of one environment collide
with each other; environments are solved together and never interact.
"""
from __future__ import annotations

import hashlib
import warnings

import numpy as np

from . import geometry as G
from .config import SimConfig
from .mesh import build_scene, check_mesh, lame, meshes_overlap
from .scene_spec import (HALFSPACE, BOX, TRIANGLE, KIND_NAMES, ActiveRegionSpec, Actuation, BodySpec,
                         ChamberSpec, ColliderSpec, RigidRegionSpec, SpringSpec, TendonSpec)


def _config_field(name, doc):
    return property(lambda self: getattr(self.cfg, name), lambda self, v: setattr(self.cfg, name, v),
                    doc=doc)


def _nonneg_mu(mu):
    if mu is None:
        return None
    mu = float(mu)
    if not (np.isfinite(mu) and mu >= 0.0):
        raise ValueError(f"mu must be a finite number >= 0, not {mu}")
    return mu


def _pixel_bodies_overlap(A, C):
    """Does the centre of a solid pixel of either body fall inside a solid pixel of the other?"""
    def centres(b):
        r, c = np.nonzero(b.mask)
        (w, h), (x0, y0) = b.size, b.center
        H, W = b.mask.shape
        return np.stack([x0 - 0.5 * w + (c + 0.5) * (w / W), y0 - 0.5 * h + (r + 0.5) * (h / H)], 1)

    def hits(pts, b):
        (w, h), (x0, y0) = b.size, b.center
        H, W = b.mask.shape
        col = np.floor((pts[:, 0] - (x0 - 0.5 * w)) / (w / W)).astype(np.int64)
        row = np.floor((pts[:, 1] - (y0 - 0.5 * h)) / (h / H)).astype(np.int64)
        on = (col >= 0) & (col < W) & (row >= 0) & (row < H)
        return bool(b.mask[row[on], col[on]].any())
    far = (abs(A.center[0] - C.center[0]) >= 0.5 * (A.size[0] + C.size[0])
           or abs(A.center[1] - C.center[1]) >= 0.5 * (A.size[1] + C.size[1]))
    return (not far) and (hits(centres(A), C) or hits(centres(C), A))


class Model:
    def __init__(self, device="cuda", config=None):
        """An empty scene. Contact parameters left unset in `config` (a SimConfig) are scaled
        to the scene at finalize(): dhat = 1/8 element, kappa = 0.004 E element / dhat,
        self_stiffness = 0.2 E, body_kappa = 0.1 E element / dhat, with E the softest soft
        body's modulus and `element` the largest element size (model.dx)."""
        if not str(device).startswith("cuda"):
            raise ValueError(f"this framework steps scenes on the GPU only (device='cuda'); "
                             f"got device={device!r}")
        self.cfg = SimConfig() if config is None else config
        self.device = device
        self._bodies: list[BodySpec] = []
        self._colliders: list[ColliderSpec] = []
        self._rigid: list[RigidRegionSpec] = []
        self.actuation = Actuation()
        self._checked_meshes = {}
        self._constrained = {}              # handle -> global ids of its prescribed nodes
        self.me = None                      # the flat scene mesh (finalize)
        self._solver = None                 # the device solver (ImplicitIntegrator.attach)
        self._node_range, self._tri_range, self._node_slice = [], [], []
        self._targets = self._x_host = self._v0 = None
        self._epoch = 0
        self.time = 0.0
        # contact between bodies: "ipc" barrier between boundary nodes and edges, or the
        # "node" penalty between boundary nodes
        self.body_contact = "ipc"
        self.body_friction = 0.0            # Coulomb coefficient between bodies (a body's own mu wins)
        self.body_kappa = None              # IPC barrier stiffness between bodies [Pa]
        self.rigid_young = 1.0e9            # stiffness a rigid region resists strain with [Pa]
        self.rigid_nu = 0.30

    dhat = _config_field("ipc_d_hat", "barrier reach [m]")
    kappa = _config_field("obstacle_stiffness", "collider barrier stiffness [Pa]")
    self_stiffness = _config_field("contact_stiffness", "'node' contact penalty [Pa]")
    mu = _config_field("obstacle_friction", "default collider friction coefficient")
    epsv = _config_field("obstacle_friction_epsv", "friction smoothing speed [m/s]")
    damping = _config_field("damping", "velocity drag [1/s]")

    # ------------------------------------------------------------------ colliders
    def add_object(self, kind, center, rx, ry=None, corner_radius=3.0e-3, mu=None):
        """A rigid signed-distance collider: kind "disc" | "ellipse" | "box" | "triangle" (or
        its integer); rx, ry half-extents (ry defaults to rx; a triangle's rx is its
        circumradius); corner_radius rounds a box's or triangle's corners. Returns its index."""
        k = KIND_NAMES[kind] if isinstance(kind, str) else int(kind)
        rx = float(rx)
        return self._new_collider(k, center, rx, rx if ry is None else float(ry),
                                  float(corner_radius) if k in (BOX, TRIANGLE) else rx, mu)

    def add_disc(self, center, radius, mu=None):
        return self.add_object("disc", center, radius, mu=mu)

    def add_plane(self, normal, origin, mu=None):
        """A wall: the half-space behind the line through `origin`, `normal` pointing into the
        free side."""
        n = np.asarray(normal, np.float64)
        size = float(np.hypot(n[0], n[1]))
        if not size > 0.0:
            raise ValueError("the plane normal must be non-zero")
        return self._new_collider(HALFSPACE, origin, float(n[0] / size), float(n[1] / size), 0.0, mu)

    def add_ground(self, height=0.0, mu=None):
        return self.add_plane((0.0, 1.0), (0.0, height), mu=mu)

    def _new_collider(self, kind, center, rx, ry, rref, mu):
        if self._solver is not None:
            raise RuntimeError("add colliders before the first simulate()")
        self._colliders.append(ColliderSpec(int(kind), float(center[0]), float(center[1]), rx, ry, rref,
                                            float(self.cfg.obstacle_friction if mu is None else mu)))
        return len(self._colliders) - 1

    def collider_records(self):
        """Every collider as the solvers read it (dicts with kind, cx, cy, rx, ry, rref, r, mu)."""
        return [c.record() for c in self._colliders]

    def set_collider_center(self, index, center):
        """Move a collider; between steps this teleports it (the move is not swept)."""
        c = self._colliders[index]
        c.cx, c.cy = float(center[0]), float(center[1])
        if self._solver is not None:
            self._solver.set_collider(index, c.record())

    # ------------------------------------------------------------------ bodies
    def _material(self, E, nu, density):
        c = self.cfg
        return (c.young if E is None else E), float(c.nu if nu is None else nu), \
            float(c.rho_mass if density is None else density)

    def _add(self, spec):
        if self.me is not None:
            raise RuntimeError("add bodies before finalize()")
        if spec.env < 0:
            raise ValueError("env_id must be >= 0")
        self._bodies.append(spec)
        return len(self._bodies) - 1

    def add_soft_body(self, mask, physical_size, center, E=None, nu=None, density=None, env_id=0, mu=None):
        """A soft body from a pixel mask ((H, W), row 0 at the bottom; two P1 triangles per
        solid pixel). E: a scalar or an (H, W) map laid out like the mask. Returns its handle."""
        if self.me is not None:
            raise RuntimeError("add bodies before finalize()")
        solid = np.asarray(mask) > 0.5
        if solid.ndim != 2 or not solid.any():
            raise ValueError("the mask must be a 2-D array with at least one solid pixel")
        if E is not None and not np.isscalar(E):
            E = np.asarray(E, np.float64)
            if E.shape != solid.shape:
                raise ValueError(f"stiffness map is {E.shape}, the mask is {solid.shape}")
            if not (E[solid] > 0).all():
                raise ValueError("stiffness map must be positive wherever the mask is solid")
        E, nu, rho = self._material(E, nu, density)
        return self._add(BodySpec(env=int(env_id), role="soft", E=float(E) if np.isscalar(E) else E,
                                  nu=nu, rho=rho, mu=_nonneg_mu(mu), mask=solid.astype(np.uint8),
                                  size=(float(physical_size[0]), float(physical_size[1])),
                                  center=(float(center[0]), float(center[1]))))

    def add_soft_mesh(self, vertices, triangles, E=None, nu=None, density=None, env_id=0, mu=None):
        """A soft body given as a triangle mesh: vertex i is its node i, triangle j its element
        j. E: a scalar or one value per triangle. Returns its handle."""
        if self.me is not None:
            raise RuntimeError("add bodies before finalize()")
        info = self._checked(vertices, triangles)
        m = len(info["triangles"])
        if E is not None and not np.isscalar(E):
            E = np.asarray(E, np.float64)
            if E.shape != (m,):
                raise ValueError(f"stiffness per triangle is {E.shape}; the mesh has {m} triangles")
            if not (np.isfinite(E) & (E > 0)).all():
                raise ValueError("stiffness must be positive and finite on every triangle")
        E, nu, rho = self._material(E, nu, density)
        V = info["vertices"]
        box = 0.5 * (V.min(0) + V.max(0))
        return self._add(BodySpec(env=int(env_id), role="soft", E=float(E) if np.isscalar(E) else E,
                                  nu=nu, rho=rho, mu=_nonneg_mu(mu), mesh=info,
                                  center=(float(box[0]), float(box[1]))))

    def _as_rigid(self, handle):
        if not self._bodies[handle].uniform:
            raise ValueError("a rigid body takes a scalar stiffness")
        self._bodies[handle].role = "affine"
        return handle

    def add_affine_body(self, mask, physical_size, center, density=None, env_id=0, E=None, nu=None, mu=None):
        """A rigid body (an affine body: x = t + A X, A held to a rotation by
        model.rigid_young) from a pixel mask. Returns its handle."""
        return self._as_rigid(self.add_soft_body(
            mask, physical_size, center, E=self.rigid_young if E is None else E,
            nu=self.rigid_nu if nu is None else nu, density=density, env_id=env_id, mu=mu))

    def add_affine_mesh(self, vertices, triangles, density=None, env_id=0, E=None, nu=None, mu=None):
        return self._as_rigid(self.add_soft_mesh(
            vertices, triangles, E=self.rigid_young if E is None else E,
            nu=self.rigid_nu if nu is None else nu, density=density, env_id=env_id, mu=mu))

    def add_kinematic_body(self, mask, physical_size, center, env_id=0, mu=None):
        """A body whose every node is prescribed (move it with set_kinematic_pose)."""
        h = self.add_soft_body(mask, physical_size, center, env_id=env_id, mu=mu)
        self._bodies[h].role = "kinematic"
        return h

    def add_kinematic_mesh(self, vertices, triangles, env_id=0, mu=None):
        h = self.add_soft_mesh(vertices, triangles, env_id=env_id, mu=mu)
        self._bodies[h].role = "kinematic"
        return h

    def _checked(self, vertices, triangles):
        """check_mesh, once per distinct mesh (a batch repeats one), with a warning when the
        elements' sizes spread over more than a factor of 10."""
        V = np.ascontiguousarray(vertices, np.float64)
        T = np.ascontiguousarray(triangles)
        digest = hashlib.sha1(V.tobytes() + T.tobytes()).hexdigest()
        key = (V.shape, T.shape, str(T.dtype), digest)
        info = self._checked_meshes.get(key)
        if info is None:
            info = check_mesh(V, T)
            side = np.sqrt(2.0 * info["area"])
            lo, hi = float(side.min()), float(side.max())
            if hi > 10.0 * lo:
                warnings.warn(f"this mesh's elements range from {lo * 1e3:.3g} to {hi * 1e3:.3g} mm; "
                              f"the iterations a step needs follow the smallest element.", stacklevel=4)
            self._checked_meshes[key] = info
        return info

    def _body_mu_per_node(self):
        """Each node's body friction coefficient, or None when no body has its own."""
        if all(b.mu is None for b in self._bodies):
            return None
        out = np.empty(self.n_nodes, np.float64)
        for h, b in enumerate(self._bodies):
            out[self._node_range[h]] = self.body_friction if b.mu is None else b.mu
        return out

    # ------------------------------------------------------------------ build
    def finalize(self):
        """Lay every body out as one flat mesh, scale the unset contact parameters, and register
        the rigid and kinematic bodies."""
        if not self._bodies:
            raise ValueError("add at least one soft body before finalize()")
        dx = self.dx
        reach = self.cfg.ipc_d_hat if self.cfg.ipc_d_hat is not None else dx / 8.0
        self.me = me = build_scene([b.scene_entry(max(2.0 * dx, 16.0 * reach)) for b in self._bodies])
        self._node_range = [np.arange(lo, hi) for lo, hi in me.body_node_range]
        self._node_slice = [slice(int(lo), int(hi)) for lo, hi in me.body_node_range]
        self._tri_range = [np.arange(lo, hi) for lo, hi in me.body_tri_range]
        for h, b in enumerate(self._bodies):          # stiffness maps scale mu and lambda together
            if not b.uniform:
                k = self._tri_range[h]
                scale = b.E if b.is_mesh else G.pixel_field_on_elements(b.E, me.tri_quad, k, b.mask.shape[0])
                me.mu[k] *= scale
                me.lam[k] *= scale
        self._check_layout()
        self._scale_contact()
        rest = np.asarray(me.rest_pos, np.float64)
        self._targets, self._x_host, self._v0 = rest.copy(), rest.copy(), np.zeros_like(rest)
        for h, b in enumerate(self._bodies):
            if b.role == "affine":
                self._register_rigid(h, self._tri_range[h])
            elif b.role == "kinematic":
                self._constrained[h] = self._node_range[h].copy()
        return self

    def _check_layout(self):
        """Bodies of one environment must start apart, and should share one element size (contact
        has one thickness, set by the largest)."""
        by_env = {}
        for h, b in enumerate(self._bodies):
            by_env.setdefault(b.env, []).append(h)
        seen = {}
        for env, members in by_env.items():
            for k, a in enumerate(members):
                for c in members[k + 1:]:
                    key = (self._shape_key(a), self._shape_key(c))
                    if key not in seen:
                        A, Cb = self._bodies[a], self._bodies[c]
                        seen[key] = (meshes_overlap(*self._rest_mesh(a), *self._rest_mesh(c))
                                     if A.is_mesh or Cb.is_mesh else _pixel_bodies_overlap(A, Cb))
                    if seen[key]:
                        raise ValueError(
                            f"bodies {a} and {c} overlap in environment {env}. Bodies of one "
                            f"environment collide with each other and must start apart; give "
                            f"the members of a batch their own environment with env_id=i.")
            sizes = [self._bodies[h].length_scale() for h in members]
            if max(sizes) > 1.25 * min(sizes):
                what = ("element sizes" if any(self._bodies[h].is_mesh for h in members)
                        else "pixel sizes")
                warnings.warn(f"environment {env} mixes {what} from {min(sizes) * 1e3:.3g} to "
                              f"{max(sizes) * 1e3:.3g} mm; contact between bodies has one thickness, "
                              f"set by the largest.", stacklevel=3)

    def _shape_key(self, h):
        b = self._bodies[h]
        if b.is_mesh:
            return ("mesh", id(b.mesh["vertices"]), id(b.mesh["triangles"]))
        return ("mask", b.mask.shape, b.mask.tobytes(), b.size, b.center)

    def _rest_mesh(self, h):
        """A body's rest mesh in its own numbering: (nodes, triangles)."""
        n0, n1 = (int(v) for v in self.me.body_node_range[h])
        return (np.asarray(self.me.rest_pos[n0:n1], np.float64),
                np.asarray(self.me.tri)[self._tri_range[h]] - n0)

    def _scale_contact(self):
        """Contact parameters the user left unset, from the scene's scales (reference
        sim_model.py:487-497): E the softest soft body, dx the largest element size."""
        c, dx = self.cfg, self.dx
        soft = [b.softest() for b in self._bodies if b.role == "soft"]
        E = min(soft) if soft else 1.0e6
        c.ipc_d_hat = dx / 8.0 if c.ipc_d_hat is None else c.ipc_d_hat
        c.obstacle_stiffness = 4.0e-3 * E * dx / c.ipc_d_hat if c.obstacle_stiffness is None \
            else c.obstacle_stiffness
        c.contact_stiffness = 0.2 * E if c.contact_stiffness is None else c.contact_stiffness
        if self.body_kappa is None:
            self.body_kappa = 0.1 * E * dx / c.ipc_d_hat

    # ------------------------------------------------------------------ rigid regions
    def add_affine_region(self, handle, region, E=None, nu=None):
        """Make part of a body one rigid piece (its interface nodes stay shared with the soft
        material). region: an (H, W) map for a pixel body; booleans per triangle or triangle
        indices for a mesh body. After finalize(), before the first step."""
        self._need_mesh()
        if self._solver is not None:
            raise RuntimeError("add rigid regions before the first simulate()")
        elems = self._region_elements(handle, region)
        mu, lam = lame(self.rigid_young if E is None else float(E), self.rigid_nu if nu is None else float(nu))
        self.me.mu[elems] = mu
        self.me.lam[elems] = lam
        return self._register_rigid(handle, elems)

    def _register_rigid(self, handle, elems):
        """A rigid piece: checked against its body's other pieces; its boundary nodes no longer
        contact one another (reference sim_model.py:454-471)."""
        elems = np.asarray(elems, np.int64)
        if any(np.isin(elems, r.elems).any() for r in self._rigid if r.body == handle):
            raise ValueError("rigid regions must not overlap")
        me = self.me
        nodes = np.unique(np.asarray(me.tri)[elems])
        env = self._bodies[handle].env
        b0, b1 = int(me.boundary_starts[env]), int(me.boundary_starts[env + 1])
        mine = np.flatnonzero(np.isin(np.asarray(me.boundary_nodes)[b0:b1], nodes))
        me.add_exclusion(env, mine)
        self._rigid.append(RigidRegionSpec(body=handle, elems=elems, nodes=nodes))
        return len(self._rigid) - 1

    @property
    def _affine(self):
        """The rigid pieces as the solvers read them: dicts with body, elems, nodes."""
        return [dict(body=r.body, elems=r.elems, nodes=r.nodes) for r in self._rigid]

    def affine_state(self, index):
        """Where rigid piece `index` is: dict(position of its rest centroid now, angle [rad],
        A (2, 2): rotation times its small strain)."""
        s = self._solver
        if s is None or s.affine is None:
            r = self._rigid[index]
            m = np.asarray(self.me.mass)[r.nodes]
            return dict(position=m @ self.me.rest_pos[r.nodes] / m.sum(), angle=0.0, A=np.eye(2))
        t, A = s.affine_state()
        a = A[index]
        return dict(position=t[index], angle=float(np.arctan2(a[1, 0] - a[0, 1], a[0, 0] + a[1, 1])), A=a)

    # ------------------------------------------------------------------ layout queries
    @property
    def n_bodies(self):
        return len(self._bodies)

    @property
    def n_envs(self):
        if self.me is not None:
            return int(self.me.n_envs)
        return 1 + max((b.env for b in self._bodies), default=-1)

    @property
    def n_nodes(self):
        return int(self.me.rest_pos.shape[0])

    @property
    def n_elements(self):
        return int(self.me.tri.shape[0])

    def body_env(self, handle):
        return self._bodies[handle].env

    def body_nodes(self, handle):
        """Global node ids of a body (rows of get_x())."""
        return self._node_range[handle].copy()

    def body_elements(self, handle):
        """Global element ids of a body."""
        return self._tri_range[handle].copy()

    def pixel_size(self, handle):
        """(width, height) of one of a pixel body's pixels [m]."""
        b = self._bodies[handle]
        if b.is_mesh:
            raise ValueError(f"body {handle} is a mesh body: it has no pixels. Its length scale is "
                             f"element_size({handle}).")
        return b.pixel()

    def element_size(self, handle):
        """A body's contact length scale [m]: its pixel width or its median boundary edge."""
        return self._bodies[handle].length_scale()

    def _cfl_length(self, handle):
        return self._bodies[handle].cfl_length()

    @property
    def dx(self):
        """The largest element size over the bodies: it sets the thickness of contact."""
        return max(b.length_scale() for b in self._bodies)

    def element_areas(self, handle):
        return np.asarray(self.me.rest_area, np.float64)[self._tri_range[handle]]

    def rest_x(self, handle):
        return np.array(self.me.rest_pos[self._node_range[handle]], np.float64)

    def to_elements(self, handle, field):
        """An (H, W) pixel field (or a mesh body's per-triangle array) on each of the body's
        elements, in body_elements order."""
        b = self._bodies[handle]
        field = np.asarray(field, np.float64)
        if b.is_mesh:
            m = len(self._tri_range[handle])
            if field.shape != (m,):
                raise ValueError(f"field is {field.shape}; mesh body {handle} has {m} triangles")
            return field.copy()
        if field.shape != b.mask.shape:
            raise ValueError(f"field is {field.shape}, the mask is {b.mask.shape}")
        return G.pixel_field_on_elements(field, self.me.tri_quad, self._tri_range[handle], b.mask.shape[0])

    def to_pixels(self, handle, values):
        """Per-element values summed back onto the body's pixel grid (to_elements transposed)."""
        b = self._bodies[handle]
        if b.is_mesh:
            raise ValueError(f"body {handle} is a mesh body: to_pixels needs a pixel body. A mesh body "
                             f"takes per-triangle arrays instead.")
        H, W = b.mask.shape
        col, row = np.divmod(np.asarray(self.me.tri_quad)[self._tri_range[handle]], H)
        grid = np.zeros(H * W)
        np.add.at(grid, row * W + col, np.asarray(values, np.float64))
        return grid.reshape(H, W)

    def rigid_motion(self, handle, translation=(0.0, 0.0), angle=0.0, about=None):
        """Rest positions of a body turned by `angle` about `about` (default: its grid's or
        bounding box's centre), then translated."""
        pivot = np.asarray(self._bodies[handle].center if about is None else about, np.float64)
        c, s = np.cos(angle), np.sin(angle)
        rel = self.rest_x(handle) - pivot
        turned = np.stack([c * rel[:, 0] - s * rel[:, 1], s * rel[:, 0] + c * rel[:, 1]], 1)
        return turned + pivot + np.asarray(translation, np.float64)

    # ------------------------------------------------------------------ kinematics and state
    def set_kinematic_constraint(self, handle, node_indices):
        """Prescribe these nodes of a body (body-local indices)."""
        self._constrained[handle] = self._node_range[handle][np.asarray(node_indices, np.int64)]

    def set_kinematic_target(self, handle, targets):
        """Where a body's nodes are to be at the end of the next step ((n_body_nodes, 2))."""
        self._targets[self._node_slice[handle]] = targets

    def set_node_targets(self, nodes, targets):
        """Targets of any nodes, by global id ((n, 2)), in one call: set_kinematic_target for
        many bodies at once. Only the prescribed nodes' targets are read by the step."""
        self._targets[np.asarray(nodes, np.int64)] = np.asarray(targets, np.float64)

    def set_kinematic_pose(self, handle, translation=(0.0, 0.0), angle=0.0, about=None):
        self.set_kinematic_target(handle, self.rigid_motion(handle, translation, angle, about))

    def all_constrained(self):
        """Global ids of every prescribed node, bodies in handle order."""
        parts = [self._constrained[h] for h in sorted(self._constrained)]
        return np.concatenate(parts) if parts else np.zeros(0, np.int64)

    def get_x(self):
        """Positions of every node ((N, 2), a copy)."""
        if self._solver is not None:
            self._x_host = self._solver.get_x()
        return self._x_host.copy()

    def get_body_x(self, handle):
        return self.get_x()[self._node_range[handle]]

    def get_body_v(self, handle):
        v = self._v0 if self._solver is None else self._solver.get_v()
        return np.array(v[self._node_range[handle]])

    def set_body_velocity(self, handle, velocity=(0.0, 0.0), omega=0.0, about=None):
        """A rigid velocity field on a body: `velocity` plus a spin `omega` about `about`
        (default: the centroid of its nodes, mass-weighted for a mesh body)."""
        nodes = self._node_range[handle]
        x = self.get_x()[nodes]
        if about is not None:
            pivot = np.asarray(about, np.float64)
        elif self._bodies[handle].is_mesh:
            m = np.asarray(self.me.mass, np.float64)[nodes]
            pivot = m @ x / m.sum()
        else:
            pivot = x.mean(0)
        r = x - pivot
        v = np.asarray(velocity, np.float64) + omega * np.stack([-r[:, 1], r[:, 0]], 1)
        self._v0[nodes] = v
        if self._solver is not None:
            self._solver.set_v_rows(nodes, v)

    def set_body_x(self, handle, x, v=None):
        """Put a body's nodes at x (and give them velocities v); its targets follow."""
        nodes = self._node_range[handle]
        x = np.asarray(x, np.float64).reshape(len(nodes), 2)
        self._x_host[nodes] = x
        self._targets[nodes] = x
        if v is not None:
            self._v0[nodes] = np.asarray(v, np.float64).reshape(len(nodes), 2)
        if self._solver is not None:
            self._solver.set_nodes(int(nodes[0]), x, None if v is None else self._v0[nodes])
            self._solver.sync_affine([i for i, r in enumerate(self._rigid) if r.body == handle])

    def set_body_pose(self, handle, translation=(0.0, 0.0), angle=0.0, about=None, velocity=(0.0, 0.0),
                      omega=0.0):
        self.set_body_x(handle, self.rigid_motion(handle, translation, angle, about))
        self.set_body_velocity(handle, velocity, omega)

    def reset(self, env_ids=None):
        """Every body (or those of the environments env_ids) back to rest at zero velocity,
        targets included; actuator commands are kept (reference sim_model.py:761-788)."""
        rest = np.asarray(self.me.rest_pos, np.float64)
        self._epoch += 1
        s = self._solver
        if env_ids is None:
            self._targets, self._x_host, self._v0 = rest.copy(), rest.copy(), np.zeros_like(rest)
            self.time = 0.0
            if s is not None:
                s.reset_all(rest)
            return
        pick = np.zeros(self.n_envs, np.int32)
        pick[np.atleast_1d(env_ids)] = 1
        rows = pick[np.asarray(self.me.node2env)] > 0
        self._targets[rows] = rest[rows]
        self._x_host[rows] = rest[rows]
        self._v0[rows] = 0.0
        if s is not None:
            s.reset_envs(pick)
            s.sync_affine([i for i, r in enumerate(self._rigid) if pick[self._bodies[r.body].env]])

    # ------------------------------------------------------------------ actuation
    # Plane strain, per unit depth: pressure in Pa, forces and tensions in N per metre of depth.
    def _need_mesh(self):
        if self.me is None:
            raise RuntimeError("call finalize() before adding actuators")

    @property
    def gravity(self):
        return self.actuation.gravity

    @gravity.setter
    def gravity(self, g):
        self.actuation.gravity = (float(g[0]), float(g[1]))
        self.actuation.commanded()

    def set_external_force(self, handle, force):
        """Nodal forces on a body ((n_body_nodes, 2) [N/m]), held until changed."""
        self._need_mesh()
        a = self.actuation
        if a.fext is None:
            a.fext = np.zeros((self.n_nodes, 2), np.float64)
        a.fext[self._node_range[handle]] = np.asarray(force, np.float64)
        a.commanded()

    # pressure ---------------------------------------------------------
    def add_pressure_chambers(self, handle, min_pixels=1):
        """Every sealed cavity of a body (a hole of a mesh body) as a pressure chamber; handles
        ordered by the cavity's lowest pixel (row, then column), or a hole's lowest node."""
        self._need_mesh()
        b = self._bodies[handle]
        if b.is_mesh:
            rest = self.me.rest_pos
            holes = [lp for lp in self._mesh_loops(handle) if lp[2] > 0]
            lowest = [rest[ea][np.argmin(rest[ea, 1])] for ea, _, _ in holes]
            order = sorted(range(len(holes)), key=lambda k: (lowest[k][1], lowest[k][0]))
            return [self._new_chamber(handle, holes[k][0], holes[k][1], holes[k][2], 1.0) for k in order]
        labels, cavities = G.enclosed_voids(b.mask)
        first = {k: tuple(np.argwhere(labels == k)[0]) for k in cavities
                 if np.count_nonzero(labels == k) >= min_pixels}
        return [self._pixel_chamber(handle, labels, k) for k in sorted(first, key=first.get)]

    def add_pressure_chamber(self, handle, at):
        """The sealed cavity (or mesh hole) containing the world point `at`, as a chamber."""
        self._need_mesh()
        b = self._bodies[handle]
        if b.is_mesh:
            rest = np.asarray(self.me.rest_pos, np.float64)
            for ea, eb, area in self._mesh_loops(handle):
                if area > 0 and G.point_in_loop(at, rest, ea, eb):
                    return self._new_chamber(handle, ea, eb, area, 1.0)
            raise ValueError(f"point {tuple(at)} is not inside a hole of this mesh body")
        px = G.pixel_of(at, b.size, b.center, b.mask.shape)
        if px is None or b.mask[px] > 0:
            raise ValueError(f"point {tuple(at)} is not inside a void of this body")
        labels, cavities = G.enclosed_voids(b.mask)
        if labels[px] not in cavities:
            raise ValueError(f"the void at {tuple(at)} is open to the outside, not a sealed cavity")
        return self._pixel_chamber(handle, labels, int(labels[px]))

    def add_ambient_pressure(self, handle):
        """A pressure acting on a body's outer surface (its outline, channels open to the
        outside included); positive squeezes the body."""
        self._need_mesh()
        b = self._bodies[handle]
        if b.is_mesh:
            outline = [lp for lp in self._mesh_loops(handle) if lp[2] < 0]
            return self._new_chamber(handle, np.concatenate([lp[0] for lp in outline]),
                                     np.concatenate([lp[1] for lp in outline]),
                                     -sum(lp[2] for lp in outline), -1.0)
        labels, cavities = G.enclosed_voids(b.mask)
        ea, eb = G.exterior_edges(b.mask, labels, cavities, self._corner_node(handle))
        cells = np.count_nonzero(b.mask) + np.count_nonzero(np.isin(labels, cavities))
        pw, ph = b.pixel()
        return self._new_chamber(handle, ea, eb, cells * pw * ph, -1.0)

    def _corner_node(self, handle):
        return lambda ix, iy: self.me.body_node_index(handle, int(ix), int(iy))

    def _pixel_chamber(self, handle, labels, label):
        b = self._bodies[handle]
        ea, eb = G.chamber_edges(b.mask, labels, label, self._corner_node(handle))
        pw, ph = b.pixel()
        return self._new_chamber(handle, ea, eb, np.count_nonzero(labels == label) * pw * ph, 1.0)

    def _mesh_loops(self, handle):
        """A mesh body's boundary loops in global node ids, checked: the outline encloses
        exactly the material and the holes."""
        n0 = int(self.me.body_node_range[handle][0])
        V, T = self._rest_mesh(handle)
        loops = [(ea + n0, eb + n0, area) for ea, eb, area in G.mesh_boundary_loops(T, V)]
        outline = -sum(area for _, _, area in loops if area < 0)
        inside = float(self.element_areas(handle).sum()) + sum(area for _, _, area in loops if area > 0)
        if abs(outline - inside) > 1e-9 * outline:
            raise RuntimeError(f"mesh body {handle}: its outline encloses {outline:.9e} m^2, its "
                               f"material and holes {inside:.9e}")
        return loops

    def _new_chamber(self, handle, ea, eb, expect, sign):
        """A pressure boundary, after checking that it closes round the expected area with the
        expected orientation."""
        rest = self.me.rest_pos
        ref = rest[np.concatenate([ea, eb])].mean(0)
        got = sign * G.loop_area(rest, ea, eb, ref)
        if not (expect > 0 and abs(got - expect) <= 1e-6 * expect):
            raise RuntimeError(f"pressure boundary encloses {got:.6e} m^2, expected {expect:.6e}")
        a = self.actuation
        return a.added(ChamberSpec(body=handle, ea=ea, eb=eb, ref=ref, area0=float(expect), sign=sign),
                       a.chambers)

    def set_pressure(self, chamber, pressure):
        """Gauge pressure in a chamber [Pa] (positive inflates)."""
        self.actuation.chambers[chamber].pressure = float(pressure)
        self.actuation.commanded()

    def chamber_area(self, chamber):
        c = self.actuation.chambers[chamber]
        return c.sign * G.loop_area(self.get_x(), c.ea, c.eb, c.ref)

    # tendons ----------------------------------------------------------
    def add_tendon(self, handle, through=None, nodes=None, stiffness=None):
        """A cable through eyelets: waypoints `through` (snapped to the nearest nodes) or body-local
        `nodes`. stiffness None: tension-controlled; k > 0: length-controlled."""
        self._need_mesh()
        if (through is None) == (nodes is None):
            raise ValueError("give exactly one of `through` (points) or `nodes` (indices)")
        ids = self._node_range[handle]
        path = (ids[np.asarray(nodes, np.int64)] if through is None
                else G.snap_path(through, self.me.rest_pos[ids], ids))
        if len(path) < 2:
            raise ValueError("a tendon needs at least two distinct eyelets")
        if stiffness is not None and not stiffness > 0:
            raise ValueError("stiffness must be positive (or None for a tension-controlled cable)")
        a = self.actuation
        return a.added(TendonSpec(body=handle, nodes=np.asarray(path, np.int64),
                                  stiffness=0.0 if stiffness is None else float(stiffness),
                                  length=G.path_length(self.me.rest_pos, path)), a.tendons)

    def set_tendon_tension(self, tendon, tension):
        t = self.actuation.tendons[tendon]
        if t.stiffness > 0:
            raise ValueError("this tendon is length-controlled; use set_tendon_length")
        t.tension = float(tension)
        self.actuation.commanded()

    def set_tendon_length(self, tendon, length):
        t = self.actuation.tendons[tendon]
        if not t.stiffness > 0:
            raise ValueError("this tendon is tension-controlled; use set_tendon_tension")
        t.length = float(length)
        self.actuation.commanded()

    def tendon_length(self, tendon):
        return G.path_length(self.get_x(), self.actuation.tendons[tendon].nodes)

    def tendon_tension(self, tendon):
        t = self.actuation.tendons[tendon]
        if t.stiffness > 0:
            return t.stiffness * max(self.tendon_length(tendon) - t.length, 0.0)
        return max(t.tension, 0.0)

    # springs and pins ---------------------------------------------------
    def _nearest_node(self, handle, point):
        ids = self._node_range[handle]
        d = self.me.rest_pos[ids] - np.asarray(point, np.float64)
        return int(ids[np.argmin((d * d).sum(1))])

    def add_spring(self, handle_a, point_a, handle_b=None, point_b=None, stiffness=None, rest_length=None):
        """A linear spring from the node of body a nearest point_a to the node of body b nearest
        point_b, or (handle_b None) to the world point point_b (default: where a snapped)."""
        self._need_mesh()
        ia = self._nearest_node(handle_a, point_a)
        if handle_b is None:
            ib = -1
            anchor = np.array(self.me.rest_pos[ia] if point_b is None else point_b, np.float64)
        else:
            if self.body_env(handle_a) != self.body_env(handle_b):
                raise ValueError("a spring cannot join bodies of different environments")
            ib = self._nearest_node(handle_b, point_a if point_b is None else point_b)
            if ib == ia:
                raise ValueError("both ends of the spring snapped to the same node")
            anchor = np.array(self.me.rest_pos[ib], np.float64)
        L = float(np.linalg.norm(self.me.rest_pos[ia] - anchor)) if rest_length is None else float(rest_length)
        if handle_b is not None and not L > 0.0:
            raise ValueError("a spring between two bodies needs a rest length: at zero it would pull two "
                             "nodes to one place while contact holds the bodies apart")
        k = float(self.rigid_young if stiffness is None else stiffness)
        if not (k > 0 and L >= 0):
            raise ValueError("spring stiffness must be positive and its rest length non-negative")
        a = self.actuation
        return a.added(SpringSpec(a=ia, b=ib, stiffness=k, length=L, anchor=anchor,
                                  env=self.body_env(handle_a)), a.springs)

    def add_pin(self, handle, point, handle_b=None, stiffness=None):
        """Pin a body's node nearest `point` to the world there (a stiff zero-length spring)."""
        if handle_b is not None:
            raise NotImplementedError("a pin between two bodies is not supported: pin to the world, "
                                      "or use add_spring with a rest length")
        return self.add_spring(handle, point, None, point, stiffness, rest_length=0.0)

    def set_spring_length(self, spring, length):
        self.actuation.springs[spring].length = float(length)
        self.actuation.commanded()

    def set_spring_anchor(self, spring, point):
        sp = self.actuation.springs[spring]
        if sp.b >= 0:
            raise ValueError("this spring joins two bodies and has no world anchor")
        sp.anchor = np.asarray(point, np.float64)
        self.actuation.commanded()

    def spring_force(self, spring):
        """(tension [N/m], current length [m]) of a spring."""
        sp = self.actuation.springs[spring]
        x = self.get_x()
        other = x[sp.b] if sp.b >= 0 else sp.anchor
        n = float(np.linalg.norm(x[sp.a] - other))
        return sp.stiffness * (n - sp.length), n

    # active material ------------------------------------------------------
    def add_active_region(self, handle, region, fiber=None):
        """Part of a body whose rest shape follows set_activation; fiber None (isotropic), an
        angle or a 2-vector."""
        self._need_mesh()
        elems = np.asarray(self._region_elements(handle, region), np.int64)
        a = self.actuation
        if any(np.isin(elems, r.elems).any() for r in a.regions if r.body == handle):
            raise ValueError("active regions must not overlap")
        return a.added(ActiveRegionSpec(body=handle, elems=elems, fiber=fiber), a.regions)

    def set_activation(self, region, stretch, transverse=1.0):
        """Rest-length ratio of an active region along its fiber (and across it)."""
        if not (stretch > 0 and transverse > 0):
            raise ValueError("stretch ratios must be positive")
        r = self.actuation.regions[region]
        r.A = G.activation_matrix(stretch, transverse, r.fiber)
        self.actuation.commanded()

    def _region_elements(self, handle, region):
        """Global element ids of part of a body: an (H, W) map of a pixel body, or booleans
        per triangle / triangle indices of a mesh body."""
        b = self._bodies[handle]
        tris = self._tri_range[handle]
        if not b.is_mesh:
            sel = np.asarray(region) > 0
            if sel.shape != b.mask.shape:
                raise ValueError(f"region is {sel.shape}, the mask is {b.mask.shape}")
            elems = G.region_elements(sel, self.me.tri_quad, tris, b.mask.shape[0])
            if not len(elems):
                raise ValueError("the region contains no solid pixel of this body")
            return elems
        m = len(tris)
        r = np.asarray(region)
        if r.dtype == bool:
            if r.shape != (m,):
                raise ValueError(f"region is a boolean array of shape {r.shape}; mesh body {handle} has "
                                 f"{m} triangles: give ({m},) booleans or triangle indices")
            local = np.flatnonzero(r)
        elif r.ndim == 1 and np.issubdtype(r.dtype, np.integer):
            local = np.unique(r)
            if len(local) != len(r) or (len(local) and not (0 <= local[0] and local[-1] < m)):
                raise ValueError(f"triangle indices must be distinct and in 0..{m - 1}")
        else:
            raise ValueError(f"region of a mesh body: ({m},) booleans or a 1-D array of triangle "
                             f"indices; got {r.dtype} of shape {r.shape}")
        if not len(local):
            raise ValueError("the region contains no triangle of this body")
        return tris[local]
