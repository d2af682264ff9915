"""Scene meshes on the host: pixel masks or given triangle meshes, laid out as one flat
multi-scene mesh for the device kernels.

The numbering here is part of the public contract (per-node targets, per-element fields and
the adjoint's per-element gradient are all addressed by it), so it follows the reference
exactly (core/mesh.py:74-109, 112-160, 380-479):

  * pixel grid corner (ix, iy) of an (nx, ny) grid is node (ny+1)*ix + iy before the nodes
    that no solid pixel touches are dropped; survivors keep their relative order;
  * each solid pixel (quad ix*ny + iy) gives the triangles [bl, tl, tr] then [bl, tr, br];
  * boundary edges are the edges used by exactly one triangle, stored (min, max), in the order
    they are first met scanning triangles in order and their edges (0,1), (1,2), (2,0);
  * boundary nodes are the sorted set of boundary-edge ends;
  * scenes (environments) are laid out in order, bodies of one scene in handle order.

The edge tables and the self-contact exclusion run in the library's host C++ (csrc/assemble.cpp,
d2f_unique_pairs / d2f_unique_triples / d2f_hop_exclusion / d2f_material_exclusion); the rest
is numpy. Nothing here touches the GPU.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field

import numpy as np


def _native():
    from . import _lib
    return _lib.lib()


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def unique_pairs(E):
    """The distinct unordered pairs among the rows of E (m, 2): (key (u, 2) sorted pairs in
    lexicographic order, first (u,) row each first occurs at, count (u,), inverse (m,)), as
    numpy.unique(numpy.sort(E, 1), axis=0, return_index, return_inverse, return_counts)."""
    E = np.ascontiguousarray(np.asarray(E, np.int64).reshape(-1, 2))
    m = len(E)
    key, inv = np.empty((m, 2), np.int64), np.empty(m, np.int64)
    first, cnt = np.empty(m, np.int64), np.empty(m, np.int64)
    u = _native().d2f_unique_pairs(_ptr(E), m, _ptr(key), _ptr(first), _ptr(cnt), _ptr(inv))
    if u < 0:
        raise ValueError("unique_pairs: bad argument")
    return key[:u], first[:u], cnt[:u], inv


def unique_triples(R):
    """The same for rows of three: (key (u, 3), first (u,), count (u,))."""
    R = np.ascontiguousarray(np.asarray(R, np.int64).reshape(-1, 3))
    m = len(R)
    key, first, cnt = np.empty((m, 3), np.int64), np.empty(m, np.int64), np.empty(m, np.int64)
    u = _native().d2f_unique_triples(_ptr(R), m, _ptr(key), _ptr(first), _ptr(cnt))
    if u < 0:
        raise ValueError("unique_triples: bad argument")
    return key[:u], first[:u], cnt[:u]


def lame(E, nu):
    """Lame moduli (mu, lambda) of Young's modulus E and Poisson ratio nu
    (reference core/mesh.py:16-19)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


@dataclass
class Mesh:
    """One flat mesh holding every body of every scene (reference core/mesh.py:22-64)."""
    n_envs: int
    rest_pos: np.ndarray          # (N, 2) float64
    tri: np.ndarray               # (T, 3) int64 global node ids
    D_m_inv: np.ndarray           # (T, 2, 2) inverse rest edge matrix, edges as columns
    rest_area: np.ndarray         # (T,)
    mu: np.ndarray                # (T,)
    lam: np.ndarray               # (T,)
    mass: np.ndarray              # (N,) lumped
    node2env: np.ndarray          # (N,)
    tri2env: np.ndarray           # (T,)
    boundary_nodes: np.ndarray    # (n_b,) global node ids
    boundary2env: np.ndarray      # (n_b,)
    boundary_edges: np.ndarray    # (E_b, 2) global node ids
    node_starts: np.ndarray       # (B+1,)
    tri_starts: np.ndarray        # (B+1,)
    boundary_starts: np.ndarray   # (B+1,)
    boundary_per_env: np.ndarray  # (B,)
    excl_flat: np.ndarray         # concatenated per-scene (n_b_e * n_b_e,) uint8
    excl_offset: np.ndarray       # (B+1,)
    excluded_pairs_per_env: list  # per scene (n_b_e, n_b_e) bool
    tri_quad: np.ndarray = None   # (T,) body-local quad index of a pixel triangle, -1 for mesh bodies
    body_env: np.ndarray = None
    body_node_range: np.ndarray = None
    body_tri_range: np.ndarray = None
    body_grid: np.ndarray = None
    grid_to_live_node_per_body: list = field(default=None)

    @property
    def sum_N(self):
        return self.rest_pos.shape[0]

    @property
    def sum_n_tri(self):
        return self.tri.shape[0]

    def body_node_index(self, body, ix, iy):
        """Global id of the node at grid corner (ix, iy) of a pixel body, -1 if pruned."""
        remap = self.grid_to_live_node_per_body[body]
        if remap is None:
            raise ValueError(f"body {body} is a mesh body: it has no pixel grid")
        ny = int(self.body_grid[body][1])
        loc = int(remap[(ny + 1) * int(ix) + int(iy)])
        return -1 if loc < 0 else int(self.body_node_range[body][0]) + loc


# ---------------------------------------------------------------------------- one body
def grid_mesh(nx, ny, Lx, Ly, center, mask):
    """Nodes and triangles of the solid pixels of an (nx, ny) grid of physical size
    (Lx, Ly) centred at `center`; returns (rest_pos, tri, tri_quad, corner->node map).
    Follows reference core/mesh.py:74-109 (the grid coordinates use the same linspace so
    that positions agree to the bit)."""
    xs = np.linspace(0.0, Lx, nx + 1) + (center[0] - 0.5 * Lx)
    ys = np.linspace(0.0, Ly, ny + 1) + (center[1] - 0.5 * Ly)
    corners = np.empty(((nx + 1) * (ny + 1), 2), np.float64)
    corners[:, 0] = np.repeat(xs, ny + 1)
    corners[:, 1] = np.tile(ys, nx + 1)

    q = np.arange(nx * ny)
    qx, qy = q // ny, q % ny
    bl = (ny + 1) * qx + qy
    tl, br = bl + 1, bl + (ny + 1)
    tr = br + 1
    tri = np.empty((2 * nx * ny, 3), np.int64)
    tri[0::2] = np.stack([bl, tl, tr], 1)
    tri[1::2] = np.stack([bl, tr, br], 1)
    quad = np.repeat(q, 2)

    if mask is not None:
        m = np.asarray(mask, bool)
        if m.shape != (nx, ny):
            raise ValueError(f"mask must be ({nx},{ny}); got {m.shape}")
        keep = m.reshape(-1)[quad]
        tri, quad = tri[keep], quad[keep]

    used = np.zeros(corners.shape[0], bool)
    used[tri.reshape(-1)] = True
    live = np.flatnonzero(used)
    remap = np.full(corners.shape[0], -1, np.int64)
    remap[live] = np.arange(live.size)
    return corners[live], remap[tri], quad, remap


def _boundary_edges(tri):
    """Edges used by exactly one triangle, (min, max), in first-met order."""
    uniq, first, cnt, _ = unique_pairs(tri[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2))
    single = cnt == 1
    order = np.argsort(first[single], kind="stable")
    return uniq[single][order]


def _hop_exclusion(tri, n_nodes, boundary, k):
    """Boundary-node pairs within k mesh-edge hops of each other (incl. each node itself),
    (reference core/mesh.py:133-153), by breadth-first search in d2f_hop_exclusion."""
    nb = boundary.size
    if nb == 0:
        return np.zeros((0, 0), bool)
    tri = np.ascontiguousarray(tri, np.int64)
    boundary = np.ascontiguousarray(boundary, np.int64)
    out = np.empty((nb, nb), np.uint8)
    rc = _native().d2f_hop_exclusion(_ptr(tri), len(tri), int(n_nodes), _ptr(boundary), nb, int(k), _ptr(out))
    if rc != 0:
        raise ValueError("hop exclusion: bad mesh indices")
    return out.astype(bool)


def material_exclusion(rest_pos, tri, boundary, radius):
    """Boundary-node pairs closer than `radius` along mesh edges at rest
    (reference core/mesh.py:163-186: the same edge graph, rest edge lengths as weights, shortest
    paths stopped at the radius), in d2f_material_exclusion."""
    n = rest_pos.shape[0]
    e, _, _, _ = unique_pairs(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]))
    pos = np.ascontiguousarray(rest_pos, np.float64)
    boundary = np.ascontiguousarray(boundary, np.int64)
    nb = len(boundary)
    out = np.zeros((nb, nb), np.uint8)
    rc = _native().d2f_material_exclusion(_ptr(pos), n, _ptr(e), len(e), _ptr(boundary), nb, float(radius),
                                          _ptr(out))
    if rc != 0:
        raise ValueError("material exclusion: bad mesh indices or radius")
    return out.astype(bool)


def finish_body(rest_pos, tri, k_excl, excl_radius=0.0):
    """Rest shape matrices, areas, boundary and self-contact exclusion of one body
    (reference core/mesh.py:112-160)."""
    p = rest_pos[tri]
    e1 = p[:, 1] - p[:, 0]
    e2 = p[:, 2] - p[:, 0]
    D_m = np.stack([e1, e2], -1)
    D_m_inv = np.linalg.inv(D_m)
    rest_area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    bedges = _boundary_edges(tri)
    boundary = np.unique(bedges.reshape(-1)).astype(np.int64)
    n = rest_pos.shape[0]
    excl = _hop_exclusion(tri, n, boundary, k_excl)
    if boundary.size and excl_radius > 0.0:
        excl |= material_exclusion(rest_pos, tri, boundary, excl_radius)
    if boundary.size:
        excl = excl | excl.T
    if bedges.size == 0:
        bedges = np.zeros((0, 2), np.int64)
    return dict(rest_pos=rest_pos, tri=tri, D_m_inv=D_m_inv, rest_area=rest_area,
                boundary=boundary, bedges=bedges, excl=excl, n_nodes=n)


# ---------------------------------------------------------------------------- mesh bodies
def _crossing_segments(Va, Ea, Vb, Eb):
    """Index pairs of segments that cross properly (each strictly separates the other's
    ends); with Va is Vb, pairs sharing a vertex are skipped."""
    from scipy.spatial import cKDTree
    A0, A1, B0, B1 = Va[Ea[:, 0]], Va[Ea[:, 1]], Vb[Eb[:, 0]], Vb[Eb[:, 1]]
    La, Lb = np.linalg.norm(A1 - A0, axis=1), np.linalg.norm(B1 - B0, axis=1)
    near = cKDTree(0.5 * (B0 + B1)).query_ball_point(0.5 * (A0 + A1),
                                                     0.5 * (La + Lb.max()) * (1 + 1e-9))
    i = np.repeat(np.arange(len(Ea)), [len(c) for c in near])
    j = (np.concatenate([np.asarray(c, np.int64) for c in near]) if len(i)
         else np.zeros(0, np.int64))
    if Va is Vb:
        keep = (i < j) & ~((Ea[i, 0] == Eb[j, 0]) | (Ea[i, 0] == Eb[j, 1]) |
                           (Ea[i, 1] == Eb[j, 0]) | (Ea[i, 1] == Eb[j, 1]))
        i, j = i[keep], j[keep]

    def cr(o, p, q):
        return (p[:, 0] - o[:, 0]) * (q[:, 1] - o[:, 1]) - (p[:, 1] - o[:, 1]) * (q[:, 0] - o[:, 0])
    a0, a1, b0, b1 = A0[i], A1[i], B0[j], B1[j]
    hit = (cr(a0, a1, b0) * cr(a0, a1, b1) < 0) & (cr(b0, b1, a0) * cr(b0, b1, a1) < 0)
    return i[hit], j[hit]


def check_mesh(vertices, triangles):
    """Validate a triangle mesh body and measure it (reference core/mesh.py:190-286).
    Raises ValueError naming the defect. Returns dict(vertices, triangles, area, h, h_cfl,
    L_max): h the median boundary-edge length, h_cfl the smallest sqrt(2 area), L_max the
    longest boundary edge."""
    V = np.asarray(vertices, np.float64)
    if V.ndim != 2 or V.shape[1] != 2 or len(V) < 3:
        raise ValueError(f"vertices must be an (n, 2) array with n >= 3; got shape {V.shape}")
    if not np.isfinite(V).all():
        raise ValueError(f"vertices {np.nonzero(~np.isfinite(V).all(1))[0][:10].tolist()} are not finite")
    T = np.asarray(triangles)
    if T.ndim != 2 or T.shape[1] != 3 or len(T) == 0:
        raise ValueError(f"triangles must be an (m, 3) array with m >= 1; got shape {T.shape}")
    if not np.issubdtype(T.dtype, np.integer):
        if not (np.issubdtype(T.dtype, np.floating) and np.all(T == np.round(T))):
            raise ValueError(f"triangles must hold integer vertex indices; got dtype {T.dtype}")
    T = T.astype(np.int64)
    if T.min() < 0 or T.max() >= len(V):
        raise ValueError(f"triangle indices run from {T.min()} to {T.max()}, but there are "
                         f"{len(V)} vertices (indices are 0-based)")
    rep = (T[:, 0] == T[:, 1]) | (T[:, 1] == T[:, 2]) | (T[:, 2] == T[:, 0])
    if rep.any():
        raise ValueError(f"triangles {np.nonzero(rep)[0][:10].tolist()} repeat a vertex")
    _, first, cnt = unique_triples(T)
    if (cnt > 1).any():
        raise ValueError(f"duplicate triangles (same three vertices), e.g. triangle {first[cnt > 1][0]}")
    used = np.zeros(len(V), bool)
    used[T.reshape(-1)] = True
    if not used.all():
        raise ValueError(f"vertices {np.nonzero(~used)[0][:10].tolist()} belong to no triangle")
    p = V[T]
    e1, e2, e3 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 2] - p[:, 1]
    twoA = np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Lmax2 = np.max([(e1 ** 2).sum(1), (e2 ** 2).sum(1), (e3 ** 2).sum(1)], 0)
    qual = twoA / Lmax2
    if (qual < 1e-8).any():
        raise ValueError(f"triangles {np.nonzero(qual < 1e-8)[0][:10].tolist()} have (almost) zero area")
    if (qual < 0.05).any():
        warnings.warn(f"{int((qual < 0.05).sum())} triangles are slivers (worst quality "
                      f"{qual.min():.3g}): they stiffen the solve and limit its step.", stacklevel=3)
    E = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    opp = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    key, _, cnt, inv = unique_pairs(E)
    if (cnt > 2).any():
        a, b = key[np.argmax(cnt > 2)]
        raise ValueError(f"edge ({a}, {b}) is shared by {cnt.max()} triangles; at most two may share one")

    def side(a, b, c):
        return ((V[b, 0] - V[a, 0]) * (V[c, 1] - V[a, 1]) - (V[b, 1] - V[a, 1]) * (V[c, 0] - V[a, 0]))
    order = np.argsort(inv, kind="stable")
    inner = np.nonzero(cnt == 2)[0]
    starts = np.concatenate([[0], np.cumsum(cnt)])[inner]
    o1, o2 = opp[order[starts]], opp[order[starts + 1]]
    s1, s2 = side(key[inner, 0], key[inner, 1], o1), side(key[inner, 0], key[inner, 1], o2)
    if (s1 * s2 >= 0).any():
        a, b = key[inner[np.argmax(s1 * s2 >= 0)]]
        raise ValueError(f"the mesh folds over at edge ({a}, {b})")
    bed = key[cnt == 1]
    if len(bed) == 0:
        raise ValueError("the mesh has no boundary")
    L = np.linalg.norm(V[bed[:, 1]] - V[bed[:, 0]], axis=1)
    from scipy.spatial import cKDTree
    tol = 1e-9 * float(np.linalg.norm(np.ptp(V, 0)))
    pairs = cKDTree(V).query_pairs(tol, output_type="ndarray")
    if len(pairs):
        raise ValueError(f"vertices {pairs[0, 0]} and {pairs[0, 1]} are at the same place; merge them")
    bnodes = np.unique(bed)
    mid = 0.5 * (V[bed[:, 0]] + V[bed[:, 1]])
    near = cKDTree(V[bnodes]).query_ball_point(mid, 0.5 * L + tol)
    ek = np.repeat(np.arange(len(bed)), [len(c) for c in near])
    if len(ek):
        cand = bnodes[np.concatenate([np.asarray(c, np.int64) for c in near])]
        keep = (cand != bed[ek, 0]) & (cand != bed[ek, 1])
        ek, cand = ek[keep], cand[keep]
        a, b = V[bed[ek, 0]], V[bed[ek, 1]]
        t = b - a
        sv = np.clip(((V[cand] - a) * t).sum(1) / np.maximum((t * t).sum(1), 1e-300), 0.0, 1.0)
        on = np.linalg.norm(V[cand] - (a + sv[:, None] * t), axis=1) < tol
        if on.any():                  # named as the reference does: the edge's first candidate
            k = ek[int(np.argmax(on))]
            c0 = cand[int(np.argmax(ek == k))]
            raise ValueError(f"vertex {c0} lies on boundary edge ({bed[k, 0]}, {bed[k, 1]}) without being "
                             f"part of it (a crack or a T-junction); merge the two sides")
    i, j = _crossing_segments(V, bed, V, bed)
    if len(i):
        raise ValueError(f"boundary edges ({bed[i[0], 0]}, {bed[i[0], 1]}) and "
                         f"({bed[j[0], 0]}, {bed[j[0], 1]}) cross: the outline intersects itself")
    return dict(vertices=V, triangles=T, area=0.5 * twoA, h=float(np.median(L)),
                h_cfl=float(np.sqrt(twoA).min()), L_max=float(L.max()))


def _triangle_pairs_overlap(P, Q, eps):
    """Do triangles P[k] and Q[k] share an area (separating axes on every edge normal)?"""
    ok = np.ones(len(P), bool)
    for S in (P, Q):
        for e in range(3):
            d = S[:, (e + 1) % 3] - S[:, e]
            n = np.stack([-d[:, 1], d[:, 0]], 1) / np.linalg.norm(d, axis=1)[:, None]
            pp, qq = np.einsum("kij,kj->ki", P, n), np.einsum("kij,kj->ki", Q, n)
            ok &= np.minimum(pp.max(1), qq.max(1)) - np.maximum(pp.min(1), qq.min(1)) > eps
    return ok


def meshes_overlap(Va, Ta, Vb, Tb):
    """Do two mesh bodies share an area of any size (reference core/mesh.py:355-377)?"""
    from scipy.spatial import cKDTree
    lo_a, hi_a, lo_b, hi_b = Va.min(0), Va.max(0), Vb.min(0), Vb.max(0)
    if (hi_a <= lo_b).any() or (hi_b <= lo_a).any():
        return False
    pa, pb = Va[Ta], Vb[Tb]
    ca, cb = pa.mean(1), pb.mean(1)
    ra = np.linalg.norm(pa - ca[:, None], axis=2).max(1)
    rb = np.linalg.norm(pb - cb[:, None], axis=2).max(1)
    eps = 1e-9 * float(np.linalg.norm(np.maximum(hi_a, hi_b) - np.minimum(lo_a, lo_b)))
    near = cKDTree(cb).query_ball_point(ca, (ra + rb.max()) * (1 + 1e-9))
    i = np.repeat(np.arange(len(Ta)), [len(n) for n in near])
    if not len(i):
        return False
    j = np.concatenate([np.asarray(n, np.int64) for n in near])
    for s in range(0, len(i), 200000):
        if _triangle_pairs_overlap(pa[i[s:s + 200000]], pb[j[s:s + 200000]], eps).any():
            return True
    return False


def mask_to_mesh(mask, physical_size, center):
    """The triangle mesh a pixel mask makes — (vertices, triangles, pixel (row, col) of each
    triangle) (reference core/mesh.py:298-310)."""
    m = (np.asarray(mask) > 0.5).T.copy()
    if m.ndim != 2 or not m.any():
        raise ValueError("the mask must be a 2-D array with at least one solid pixel")
    nx, ny = m.shape
    rest_pos, tri, tri_quad, _ = grid_mesh(nx, ny, float(physical_size[0]), float(physical_size[1]),
                                           (float(center[0]), float(center[1])), m)
    return rest_pos, tri, np.stack([tri_quad % ny, tri_quad // ny], 1)


# ---------------------------------------------------------------------------- the scene
def build_scene(bodies, k_excl=2):
    """Lay every body of every scene out as one flat Mesh (reference core/mesh.py:380-479).

    bodies: dicts with either mask_xy (nx, ny) bool + size + center, or vertices +
    triangles (+ excl_radius, checked); plus young, nu, rho, env."""
    nbod = len(bodies)
    envs = np.array([int(b["env"]) for b in bodies])
    n_env = int(envs.max()) + 1
    if set(envs.tolist()) != set(range(n_env)):
        raise ValueError(f"environment ids must be 0..{n_env - 1} with none skipped; "
                         f"got {sorted(set(envs.tolist()))}")
    order = np.argsort(envs, kind="stable")
    per, cache = {}, {}
    for h in order:
        b = bodies[h]
        if "mask_xy" in b:
            m = np.asarray(b["mask_xy"], bool)
            key = (m.shape, m.tobytes(), tuple(b["size"]), tuple(b["center"]))
            if key not in cache:
                pos, tri, quad, remap = grid_mesh(m.shape[0], m.shape[1], b["size"][0], b["size"][1],
                                                  b["center"], m)
                out = finish_body(pos, tri, k_excl)
                out.update(grid_to_live=remap, tri_quad=quad)
                cache[key] = out
        else:
            V = np.ascontiguousarray(b["vertices"], np.float64)
            T = np.ascontiguousarray(b["triangles"], np.int64)
            r = float(b.get("excl_radius", 0.0))
            key = ("mesh", V.shape, V.tobytes(), T.shape, T.tobytes(), r)
            if key not in cache:
                if not b.get("checked"):
                    check_mesh(V, T)
                out = finish_body(V, T, k_excl, r)
                out.update(grid_to_live=None, tri_quad=np.full(len(T), -1, np.int64))
                cache[key] = out
        per[h] = cache[key]

    node_rng = np.zeros((nbod, 2), np.int64)
    tri_rng = np.zeros((nbod, 2), np.int64)
    n0 = t0 = 0
    for h in order:
        node_rng[h] = (n0, n0 + per[h]["n_nodes"])
        n0 = node_rng[h][1]
        tri_rng[h] = (t0, t0 + len(per[h]["rest_area"]))
        t0 = tri_rng[h][1]

    def cat(key):
        return np.concatenate([per[h][key] for h in order], 0)
    rest_pos, D_m_inv, rest_area, tri_quad = cat("rest_pos"), cat("D_m_inv"), cat("rest_area"), cat("tri_quad")
    tri = np.concatenate([per[h]["tri"] + node_rng[h][0] for h in order], 0)
    boundary = np.concatenate([per[h]["boundary"] + node_rng[h][0] for h in order], 0)
    be = [per[h]["bedges"] + node_rng[h][0] for h in order if len(per[h]["bedges"])]
    boundary_edges = np.concatenate(be, 0) if be else np.zeros((0, 2), np.int64)

    def per_item(count):
        return np.concatenate([np.full(count(h), envs[h]) for h in order])
    node2env = per_item(lambda h: per[h]["n_nodes"])
    tri2env = per_item(lambda h: len(per[h]["rest_area"]))
    boundary2env = per_item(lambda h: len(per[h]["boundary"]))

    def starts(a):
        return np.concatenate([[0], np.cumsum(a)])
    members = [[] for _ in range(n_env)]
    for h in order:
        members[envs[h]].append(h)
    b_pe = np.array([sum(len(per[h]["boundary"]) for h in mem) for mem in members])
    excls = []
    for e, mem in enumerate(members):
        ex = np.zeros((b_pe[e], b_pe[e]), bool)
        o = 0
        for h in mem:
            k = len(per[h]["boundary"])
            ex[o:o + k, o:o + k] = per[h]["excl"]
            o += k
        excls.append(ex)
    excl_flat = (np.concatenate([e.reshape(-1).astype(np.uint8) for e in excls])
                 if excls else np.zeros(0, np.uint8))
    excl_offset = starts([e.size for e in excls]).astype(np.int64)

    mu = np.zeros(len(rest_area))
    lam = np.zeros(len(rest_area))
    mass = np.zeros(rest_pos.shape[0])
    for h in order:
        b = bodies[h]
        ts = slice(*tri_rng[h])
        mu[ts], lam[ts] = lame(b["young"], b["nu"])
        np.add.at(mass, tri[ts].reshape(-1), np.repeat(b["rho"] * rest_area[ts] / 3.0, 3))
    mass = np.clip(mass, 1e-30, None)

    grid = np.array([np.asarray(bodies[h]["mask_xy"]).shape if "mask_xy" in bodies[h] else (0, 0)
                     for h in range(nbod)])
    return Mesh(n_envs=n_env, rest_pos=rest_pos, tri=tri, D_m_inv=D_m_inv, rest_area=rest_area,
                mu=mu, lam=lam, mass=mass, node2env=node2env, tri2env=tri2env,
                boundary_nodes=boundary, boundary2env=boundary2env, boundary_edges=boundary_edges,
                node_starts=starts([sum(per[h]["n_nodes"] for h in mem) for mem in members]),
                tri_starts=starts([sum(len(per[h]["rest_area"]) for h in mem) for mem in members]),
                boundary_starts=starts(b_pe), boundary_per_env=b_pe,
                excl_flat=excl_flat, excl_offset=excl_offset, excluded_pairs_per_env=excls,
                tri_quad=tri_quad, body_env=envs, body_node_range=node_rng, body_tri_range=tri_rng,
                body_grid=grid, grid_to_live_node_per_body=[per[h]["grid_to_live"] for h in range(nbod)])
