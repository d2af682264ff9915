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
d2f_unique_pairs / d2f_unique_triples / d2f_exclusion_csr); the rest is numpy. The exclusion
is kept as sorted lists per boundary node (CSR), not the reference's dense n_b_e x n_b_e
matrices; Mesh builds those views on demand. Nothing here touches the GPU.
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
    # self-contact exclusion as sorted lists (CSR over the boundary nodes in layout order):
    # boundary node b excludes the scene-local boundary indices excl_idx[excl_ptr[b]:excl_ptr[b+1]]
    # (the properties excl_ptr, excl_idx; `excl` holds (ptr (n_b + 1,) int64, idx int32) as of
    # the last exclusions applied)
    excl: tuple
    tri_quad: np.ndarray = None   # (T,) body-local quad index of a pixel triangle, -1 for mesh bodies
    body_env: np.ndarray = None
    body_node_range: np.ndarray = None
    body_tri_range: np.ndarray = None
    body_grid: np.ndarray = None
    grid_to_live_node_per_body: list = field(default=None)
    # exclusions added since `excl` was built: (env, local boundary indices), applied together
    # the next time the lists are read
    _excl_pending: list = field(default_factory=list, repr=False)

    @property
    def excl_ptr(self):
        self._apply_exclusions()
        return self.excl[0]

    @property
    def excl_idx(self):
        self._apply_exclusions()
        return self.excl[1]

    @property
    def sum_N(self):
        return self.rest_pos.shape[0]

    @property
    def sum_n_tri(self):
        return self.tri.shape[0]

    # ---- the reference's dense views of the exclusion (core/mesh.py:452-459), built on demand
    @property
    def excl_offset(self):
        """(B+1,) offsets of each scene's n_b_e x n_b_e block in excl_flat."""
        n = np.asarray(self.boundary_per_env, np.int64)
        return np.concatenate([[0], np.cumsum(n * n)]).astype(np.int64)

    def excluded_pairs(self, env):
        """Scene env's exclusion as an (n_b_e, n_b_e) bool matrix (a fresh array)."""
        b0, b1 = int(self.boundary_starts[env]), int(self.boundary_starts[env + 1])
        n = b1 - b0
        out = np.zeros((n, n), bool)
        p = self.excl_ptr
        rows = np.repeat(np.arange(n), np.diff(p[b0:b1 + 1]))
        out[rows, self.excl_idx[p[b0]:p[b1]]] = True
        return out

    @property
    def excluded_pairs_per_env(self):
        return [self.excluded_pairs(e) for e in range(self.n_envs)]

    @property
    def excl_flat(self):
        """Every scene's block, flattened and concatenated, uint8 (reference layout)."""
        blocks = [self.excluded_pairs(e).reshape(-1).astype(np.uint8) for e in range(self.n_envs)]
        return np.concatenate(blocks) if blocks else np.zeros(0, np.uint8)

    def add_exclusion(self, env, local):
        """Exclude every pair of scene env's boundary nodes with local indices `local`
        (each with itself too). Recorded here and applied, with every other exclusion added
        since, the next time excl_ptr / excl_idx are read: one rebuild of the lists instead of
        one per call."""
        local = np.unique(np.asarray(local, np.int64))
        if local.size:
            self._excl_pending.append((int(env), local))

    def _apply_exclusions(self):
        if not self._excl_pending:
            return
        groups = {}
        for env, local in self._excl_pending:
            groups.setdefault(env, []).append(local)
        self._excl_pending = []
        p, idx = self.excl
        counts = np.diff(p)
        pieces, at = [], 0
        for env in sorted(groups):
            b0, b1 = int(self.boundary_starts[env]), int(self.boundary_starts[env + 1])
            rows = [idx[p[b]:p[b + 1]] for b in range(b0, b1)]
            for local in groups[env]:
                for j in local:
                    rows[j] = np.union1d(rows[j], local).astype(np.int32)
            pieces.append(idx[p[at]:p[b0]])
            pieces.extend(rows)
            counts[b0:b1] = [len(r) for r in rows]
            at = b1
        pieces.append(idx[p[at]:])
        ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.excl = (ptr, np.concatenate(pieces).astype(np.int32))

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


def exclusion_csr(tri, n_nodes, boundary, k, rest_pos=None, radius=0.0):
    """Self-contact exclusion of one body as sorted lists over its boundary nodes: pairs within
    k mesh-edge hops (each node with itself; reference core/mesh.py:133-153), and with
    radius > 0 pairs closer than radius along mesh edges at rest (reference
    core/mesh.py:163-186), made symmetric (core/mesh.py:157-158). Returns (ptr (nb+1,) int64,
    idx int32): node s excludes boundary indices idx[ptr[s]:ptr[s+1]]. Built in
    d2f_exclusion_csr (breadth-first search and Dijkstra per boundary node) without the
    nb x nb matrix of the reference."""
    nb = int(np.asarray(boundary).size)
    ptr = np.zeros(nb + 1, np.int64)
    if nb == 0:
        return ptr, np.zeros(0, np.int32)
    tri = np.ascontiguousarray(tri, np.int64)
    boundary = np.ascontiguousarray(boundary, np.int64)
    pos = None if rest_pos is None else np.ascontiguousarray(rest_pos, np.float64)
    if radius > 0.0 and pos is None:
        raise ValueError("material exclusion needs the rest positions")
    cap = 16 * nb
    while True:
        idx = np.empty(cap, np.int32)
        nnz = _native().d2f_exclusion_csr(_ptr(tri), len(tri), int(n_nodes), None if pos is None else _ptr(pos),
                                          _ptr(boundary), nb, int(k), float(radius), _ptr(ptr), _ptr(idx), cap)
        if nnz < 0:
            raise ValueError("self-contact exclusion: bad mesh indices or radius")
        if nnz <= cap:
            return ptr, idx[:nnz].copy()
        cap = int(nnz)


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
    ptr, idx = exclusion_csr(tri, n, boundary, k_excl, rest_pos, excl_radius)
    if bedges.size == 0:
        bedges = np.zeros((0, 2), np.int64)
    return dict(rest_pos=rest_pos, tri=tri, D_m_inv=D_m_inv, rest_area=rest_area,
                boundary=boundary, bedges=bedges, excl_ptr=ptr, excl_idx=idx, n_nodes=n)


# ---------------------------------------------------------------------------- mesh bodies
def _sweep_pairs(lo_a, hi_a, lo_b, hi_b):
    """Index pairs (i, j) of boxes [lo_a[i], hi_a[i]] and [lo_b[j], hi_b[j]] (2-D) whose
    closures overlap, found by sorting the boxes on x and sweeping: box i meets the b boxes
    whose x-interval starts before hi_a[i] among those starting after lo_a[i] - (widest b)."""
    if len(lo_a) == 0 or len(lo_b) == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    ob = np.argsort(lo_b[:, 0], kind="stable")
    xs = lo_b[ob, 0]
    wmax = float((hi_b[:, 0] - lo_b[:, 0]).max())
    first = np.searchsorted(xs, lo_a[:, 0] - wmax, side="left")
    last = np.searchsorted(xs, hi_a[:, 0], side="right")
    cnt = np.maximum(last - first, 0)
    i = np.repeat(np.arange(len(lo_a)), cnt)
    if len(i) == 0:
        return i, i
    off = np.arange(len(i)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    j = ob[np.repeat(first, cnt) + off]
    keep = ((lo_a[i] <= hi_b[j]) & (lo_b[j] <= hi_a[i])).all(1)
    return i[keep], j[keep]


def _orient(o, p, q):
    """z of (p - o) x (q - o), row-wise."""
    return (p[:, 0] - o[:, 0]) * (q[:, 1] - o[:, 1]) - (q[:, 0] - o[:, 0]) * (p[:, 1] - o[:, 1])


def _crossing_segments(Va, Ea, Vb, Eb):
    """Index pairs (i, j) of segments Ea[i] of Va and Eb[j] of Vb that cross properly: each
    strictly separates the other's two ends (touching, collinear overlap and shared ends do
    not count). With Va is Vb, pairs that share a vertex are skipped. Candidates are pairs of
    overlapping bounding boxes (_sweep_pairs)."""
    A0, A1, B0, B1 = Va[Ea[:, 0]], Va[Ea[:, 1]], Vb[Eb[:, 0]], Vb[Eb[:, 1]]
    i, j = _sweep_pairs(np.minimum(A0, A1), np.maximum(A0, A1), np.minimum(B0, B1), np.maximum(B0, B1))
    if Va is Vb:
        share = (Ea[i][:, :, None] == Eb[j][:, None, :]).any((1, 2))
        keep = (i < j) & ~share
        i, j = i[keep], j[keep]
    a0, a1, b0, b1 = A0[i], A1[i], B0[j], B1[j]
    hit = (np.sign(_orient(a0, a1, b0)) * np.sign(_orient(a0, a1, b1)) < 0) & \
          (np.sign(_orient(b0, b1, a0)) * np.sign(_orient(b0, b1, a1)) < 0)
    return i[hit], j[hit]


def _check_arrays(vertices, triangles):
    """The mesh as (V float64 (n, 2), T int64 (m, 3)); ValueError for a malformed one."""
    V = np.asarray(vertices, np.float64)
    if V.ndim != 2 or V.shape[1] != 2 or len(V) < 3:
        raise ValueError(f"vertices must be an (n, 2) array with n >= 3; got shape {V.shape}")
    bad = np.flatnonzero(~np.isfinite(V).all(1))
    if len(bad):
        raise ValueError(f"vertices {bad[:10].tolist()} are not finite")
    T = np.asarray(triangles)
    if T.ndim != 2 or T.shape[1] != 3 or len(T) == 0:
        raise ValueError(f"triangles must be an (m, 3) array with m >= 1; got shape {T.shape}")
    integral = np.issubdtype(T.dtype, np.integer) or (
        np.issubdtype(T.dtype, np.floating) and bool(np.all(np.floor(T) == T)))
    if not integral:
        raise ValueError(f"triangles must hold integer vertex indices; got dtype {T.dtype}")
    T = T.astype(np.int64)
    lo, hi = int(T.min()), int(T.max())
    if lo < 0 or hi >= len(V):
        raise ValueError(f"triangle indices run from {lo} to {hi}, but there are {len(V)} vertices "
                         f"(indices are 0-based)")
    return V, T


def _check_triangles(V, T):
    """Each triangle's own defects: a repeated vertex, a duplicate, unused vertices, (almost)
    zero area; slivers warn. Returns twice each triangle's area."""
    same = (T == np.roll(T, 1, axis=1)).any(1)
    if same.any():
        raise ValueError(f"triangles {np.flatnonzero(same)[:10].tolist()} repeat a vertex")
    _, first, cnt = unique_triples(T)
    if (cnt > 1).any():
        raise ValueError(f"duplicate triangles (same three vertices), e.g. triangle {first[cnt > 1][0]}")
    unused = np.setdiff1d(np.arange(len(V)), T)
    if len(unused):
        raise ValueError(f"vertices {unused[:10].tolist()} belong to no triangle; remove them "
                         f"(vertex i is node i of the body, so none may be left over)")
    P = V[T]
    sides = P[:, [1, 2, 2]] - P[:, [0, 0, 1]]                        # e1, e2, e3
    twoA = np.abs(sides[:, 0, 0] * sides[:, 1, 1] - sides[:, 1, 0] * sides[:, 0, 1])
    quality = twoA / (sides ** 2).sum(2).max(1)                     # 0.87 equilateral, 0.5 a pixel triangle
    flat = np.flatnonzero(quality < 1e-8)
    if len(flat):
        raise ValueError(f"triangles {flat[:10].tolist()} have (almost) zero area")
    thin = quality < 0.05
    if thin.any():
        warnings.warn(f"{int(thin.sum())} triangles are slivers (2 x area / longest edge^2 below "
                      f"0.05, worst {quality.min():.3g}): they stiffen the solve and limit its "
                      f"step. Remesh with a quality constraint.", stacklevel=4)
    return twoA


def _check_edges(V, T):
    """Every edge is shared by at most two triangles, and two triangles sharing one lie on
    opposite sides of it. A triangle's third vertex lies left of an edge it traverses when the
    triangle's own vertex order is counter-clockwise, so two triangles are on opposite sides of
    their shared edge exactly when their orientations times their traversal directions agree.
    Returns the boundary edges (sorted pairs, in lexicographic order)."""
    half = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    key, _, cnt, inv = unique_pairs(half)
    if (cnt > 2).any():
        a, b = key[np.argmax(cnt > 2)]
        raise ValueError(f"edge ({a}, {b}) is shared by {cnt.max()} triangles; at most two may share one")
    P = V[T]
    ccw = np.sign(_orient(P[:, 0], P[:, 1], P[:, 2]))
    sense = np.tile(ccw, 3) * np.where(half[:, 0] < half[:, 1], 1.0, -1.0)
    inner = cnt == 2
    by_edge = np.argsort(inv, kind="stable")
    head = np.concatenate([[0], np.cumsum(cnt)[:-1]])[inner]
    s1, s2 = sense[by_edge[head]], sense[by_edge[head + 1]]
    fold = np.flatnonzero(s1 == s2)
    if len(fold):
        a, b = key[np.flatnonzero(inner)[fold[0]]]
        raise ValueError(f"the mesh folds over at edge ({a}, {b}): its two triangles lie on the "
                         f"same side of it")
    bed = key[cnt == 1]
    if len(bed) == 0:
        raise ValueError("the mesh has no boundary")
    return bed


def _check_outline(V, bed):
    """No two vertices at one place, no vertex on a boundary edge it is not part of (a crack or
    a T-junction), no two boundary edges crossing."""
    from scipy.spatial import cKDTree
    tol = 1e-9 * float(np.linalg.norm(V.max(0) - V.min(0)))
    dup = cKDTree(V).query_pairs(tol, output_type="ndarray")
    if len(dup):
        raise ValueError(f"vertices {dup[0, 0]} and {dup[0, 1]} are at the same place; merge them")
    bnodes = np.unique(bed)
    a, b = V[bed[:, 0]], V[bed[:, 1]]
    k, n = _sweep_pairs(np.minimum(a, b) - tol, np.maximum(a, b) + tol, V[bnodes], V[bnodes])
    c = bnodes[n]
    k, c = k[(c != bed[k, 0]) & (c != bed[k, 1])], c[(c != bed[k, 0]) & (c != bed[k, 1])]
    if len(k):
        d = b[k] - a[k]
        w = V[c] - a[k]
        t = np.clip((w * d).sum(1) / np.maximum((d * d).sum(1), 1e-300), 0.0, 1.0)
        on = np.hypot(w[:, 0] - t * d[:, 0], w[:, 1] - t * d[:, 1]) < tol
        if on.any():
            e = k[on].min()
            v = c[on][k[on] == e].min()
            raise ValueError(f"vertex {v} lies on boundary edge ({bed[e, 0]}, {bed[e, 1]}) without being "
                             f"part of it (a crack or a T-junction); merge the two sides")
    i, j = _crossing_segments(V, bed, V, bed)
    if len(i):
        first = np.lexsort((j, i))[0]
        i, j = i[first], j[first]
        raise ValueError(f"boundary edges ({bed[i, 0]}, {bed[i, 1]}) and ({bed[j, 0]}, {bed[j, 1]}) "
                         f"cross: the outline intersects itself")


def check_mesh(vertices, triangles):
    """Validate a triangle mesh body and measure it (reference core/mesh.py:190-286: the same
    defects are refused, with ValueError naming them). Either winding is fine, mixed included.

    Returns dict(vertices (n, 2) float64, triangles (m, 3) int64, area (m,), h, h_cfl, L_max):
    h the median boundary-edge length (the body's contact length scale), h_cfl the smallest
    sqrt(2 x area) (the size that sets the iterations a step needs), L_max the longest
    boundary edge."""
    V, T = _check_arrays(vertices, triangles)
    twoA = _check_triangles(V, T)
    bed = _check_edges(V, T)
    d = V[bed[:, 1]] - V[bed[:, 0]]
    L = np.sqrt((d * d).sum(1))
    _check_outline(V, bed)
    return dict(vertices=V, triangles=T, area=0.5 * twoA, h=float(np.median(L)),
                h_cfl=float(np.sqrt(twoA.min())), L_max=float(L.max()))


def _triangles_overlap(P, Q, eps):
    """Do triangles P[k] and Q[k] ((K, 3, 2) each) share an area? Separating-axis test: two
    convex polygons overlap in area exactly when, on the normal of every edge of either, their
    projections overlap by more than eps (touching along an edge or at a point is no overlap)."""
    both = np.concatenate([P, Q], 1)                                 # (K, 6, 2)
    edges = np.concatenate([np.roll(P, -1, 1) - P, np.roll(Q, -1, 1) - Q], 1)   # (K, 6, 2)
    normal = np.stack([-edges[..., 1], edges[..., 0]], -1)
    normal /= np.linalg.norm(normal, axis=-1, keepdims=True)
    proj = np.einsum("kvd,kad->kav", both, normal)                   # (K, axis, vertex)
    gap = np.minimum(proj[..., :3].max(-1), proj[..., 3:].max(-1)) - \
        np.maximum(proj[..., :3].min(-1), proj[..., 3:].min(-1))
    return (gap > eps).all(1)


def meshes_overlap(Va, Ta, Vb, Tb):
    """Do two mesh bodies share an area of any size? Exact for any alignment: every pair of
    triangles whose bounding boxes meet is tested (reference core/mesh.py:355-377)."""
    lo_a, hi_a, lo_b, hi_b = Va.min(0), Va.max(0), Vb.min(0), Vb.max(0)
    if (hi_a <= lo_b).any() or (hi_b <= lo_a).any():
        return False
    pa, pb = Va[Ta], Vb[Tb]
    i, j = _sweep_pairs(pa.min(1), pa.max(1), pb.min(1), pb.max(1))
    eps = 1e-9 * float(np.linalg.norm(np.maximum(hi_a, hi_b) - np.minimum(lo_a, lo_b)))
    for s in range(0, len(i), 1 << 17):                              # bounded memory
        if _triangles_overlap(pa[i[s:s + (1 << 17)]], pb[j[s:s + (1 << 17)]], eps).any():
            return True
    return False


def mask_to_mesh(mask, physical_size, center):
    """The mesh a pixel body of this mask is made of — the same nodes and triangles in the same
    order — as (vertices (n, 2), triangles (m, 3), pixel (m, 2)): pixel[j] is the (row, col) of
    the mask pixel triangle j comes from (reference core/mesh.py:298-310)."""
    solid = np.asarray(mask) > 0.5
    if solid.ndim != 2 or not solid.any():
        raise ValueError("the mask must be a 2-D array with at least one solid pixel")
    H, W = solid.shape
    V, T, quad, _ = grid_mesh(W, H, float(physical_size[0]), float(physical_size[1]),
                              (float(center[0]), float(center[1])), np.ascontiguousarray(solid.T))
    col, row = np.divmod(quad, H)
    return V, T, np.stack([row, col], 1)


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
    # exclusion lists, block-diagonal per scene: body h's local boundary indices are shifted by
    # the boundary nodes of the bodies before it in its scene
    counts, cols = [], []
    for mem in members:
        o = 0
        for h in mem:
            counts.append(np.diff(per[h]["excl_ptr"]))
            cols.append(per[h]["excl_idx"] + o)
            o += len(per[h]["boundary"])
    excl_ptr = starts(np.concatenate(counts) if counts else np.zeros(0, np.int64)).astype(np.int64)
    excl_idx = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)

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
                excl=(excl_ptr, excl_idx),
                tri_quad=tri_quad, body_env=envs, body_node_range=node_rng, body_tri_range=tri_rng,
                body_grid=grid, grid_to_live_node_per_body=[per[h]["grid_to_live"] for h in range(nbod)])
