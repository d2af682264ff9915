"""What a Model is made of, as plain records: bodies, colliders, rigid and active regions and
the actuators, plus the arrays the solvers are built from.

The Model (model.py) is the user-facing scene API; these records are what it keeps. Both the
device solver and the CPU oracle read the actuators through Actuation's *_layout methods, so
the two are always built from the same arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# collider kinds (reference core/objects.py:22-24)
NONE, DISC, ELLIPSE, BOX, TRIANGLE, HALFSPACE = -1, 0, 1, 2, 3, 4
KIND_NAMES = {"disc": DISC, "ellipse": ELLIPSE, "box": BOX, "square": BOX, "triangle": TRIANGLE}


@dataclass
class ColliderSpec:
    """A fixed analytic collider: kind, centre, half-extents, corner rounding (rref; a disc's
    and an ellipse's is its rx), Coulomb coefficient. A half-space keeps its unit normal in
    (rx, ry)."""
    kind: int
    cx: float
    cy: float
    rx: float
    ry: float
    rref: float
    mu: float

    def record(self):
        """The collider as the solvers read it: a dict with r = rx."""
        return dict(kind=self.kind, cx=self.cx, cy=self.cy, rx=self.rx, ry=self.ry, rref=self.rref,
                    r=self.rx, mu=self.mu)


@dataclass
class BodySpec:
    """One body as added. A pixel body has mask / size / center (of its grid); a mesh body has
    its checked mesh (vertices, triangles, h, h_cfl, L_max) and center = its bounding box's.
    role: "soft", "affine" (rigid as a whole) or "kinematic" (every node prescribed)."""
    env: int
    role: str
    E: object                      # float, or (H, W) map / (m,) per triangle
    nu: float
    rho: float
    mu: Optional[float]
    center: tuple
    mask: Optional[np.ndarray] = None
    size: Optional[tuple] = None
    mesh: Optional[dict] = None

    @property
    def is_mesh(self):
        return self.mesh is not None

    @property
    def uniform(self):
        return np.isscalar(self.E)

    def pixel(self):
        """(width, height) of one pixel."""
        H, W = self.mask.shape
        return self.size[0] / W, self.size[1] / H

    def length_scale(self):
        """Contact length scale: the pixel width, or a mesh body's median boundary edge."""
        return float(self.mesh["h"]) if self.is_mesh else self.pixel()[0]

    def cfl_length(self):
        """The element size a wave crosses in one element: the smaller pixel side, or a mesh
        body's smallest sqrt(2 x triangle area)."""
        return float(self.mesh["h_cfl"]) if self.is_mesh else min(self.pixel())

    def softest(self):
        """Smallest modulus over its material."""
        if self.uniform:
            return float(self.E)
        return float(self.E.min()) if self.is_mesh else float(self.E[self.mask > 0].min())

    def scene_entry(self, excl_radius):
        """The body as build_scene takes it; a stiffness map is applied after the build, so the
        mesh is built at unit modulus then."""
        geo = (dict(vertices=self.mesh["vertices"], triangles=self.mesh["triangles"],
                    excl_radius=excl_radius, checked=True) if self.is_mesh else
               dict(mask_xy=np.ascontiguousarray(self.mask.T), size=self.size, center=self.center))
        return dict(geo, young=float(self.E) if self.uniform else 1.0, nu=self.nu, rho=self.rho,
                    env=self.env)


@dataclass
class RigidRegionSpec:
    """Elements of one body that move as one affine body, and their nodes."""
    body: int
    elems: np.ndarray
    nodes: np.ndarray


@dataclass
class ChamberSpec:
    """A pressure boundary: directed edges (ea -> eb) with the fluid on the left, the point the
    enclosed area is measured about, the area at rest, sign +1 for a cavity (counter-clockwise)
    or -1 for the surroundings of the outline (clockwise), and its gauge pressure."""
    body: int
    ea: np.ndarray
    eb: np.ndarray
    ref: np.ndarray
    area0: float
    sign: float
    pressure: float = 0.0


@dataclass
class TendonSpec:
    """A cable through eyelet nodes: stiffness 0 = tension-controlled (tension), else
    length-controlled with free length `length`."""
    body: int
    nodes: np.ndarray
    stiffness: float
    length: float
    tension: float = 0.0


@dataclass
class SpringSpec:
    """A linear spring between nodes a and b, or from a to the world point `anchor` (b = -1)."""
    a: int
    b: int
    stiffness: float
    length: float
    anchor: np.ndarray
    env: int


@dataclass
class ActiveRegionSpec:
    """Elements whose rest shape follows an activation A (2 x 2); fiber as given."""
    body: int
    elems: np.ndarray
    fiber: object
    A: np.ndarray = field(default_factory=lambda: np.eye(2))


class Actuation:
    """The actuators of a scene and their commands. `structure` changes when an actuator is
    added (the solvers rebuild their terms), `commands` when a command changes (the solvers
    push new values)."""

    def __init__(self):
        self.chambers: list[ChamberSpec] = []
        self.tendons: list[TendonSpec] = []
        self.springs: list[SpringSpec] = []
        self.regions: list[ActiveRegionSpec] = []
        self.gravity = (0.0, 0.0)
        self.fext = None                      # (N, 2) applied nodal force, or None
        self.structure = 0
        self.commands = 0

    def added(self, item, into):
        into.append(item)
        self.structure += 1
        return len(into) - 1

    def commanded(self):
        self.commands += 1

    # ---- the arrays the solvers are built from
    def pressure_layout(self, env_of_body):
        """Every chamber's edges concatenated: ea, eb, the chamber of each edge, each chamber's
        environment and reference point."""
        C = self.chambers
        return dict(ea=np.concatenate([c.ea for c in C]), eb=np.concatenate([c.eb for c in C]),
                    chamber=np.concatenate([np.full(len(c.ea), i) for i, c in enumerate(C)]),
                    env=[env_of_body(c.body) for c in C], ref=[c.ref for c in C])

    def tendon_layout(self, env_of_body):
        T = self.tendons
        return dict(paths=[t.nodes for t in T], env=[env_of_body(t.body) for t in T],
                    stiffness=[t.stiffness for t in T], length=[t.length for t in T])

    def spring_layout(self):
        S = self.springs
        return dict(a=[s.a for s in S], b=[s.b for s in S], stiffness=[s.stiffness for s in S],
                    env=[s.env for s in S])

    def region_layout(self):
        R = self.regions
        return dict(elems=np.concatenate([r.elems for r in R]),
                    region=np.concatenate([np.full(len(r.elems), i) for i, r in enumerate(R)]),
                    n=len(R))

    def external_force(self, mass):
        """(N, 2) nodal force of gravity plus the applied forces, or None when there is none."""
        if self.fext is None and self.gravity == (0.0, 0.0):
            return None
        f = np.zeros((len(mass), 2)) if self.fext is None else np.array(self.fext, np.float64)
        f += np.asarray(mass, np.float64)[:, None] * np.asarray(self.gravity, np.float64)[None, :]
        return f
