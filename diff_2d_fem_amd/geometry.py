"""Host geometry behind the actuators: which mesh entities a pressure chamber, a tendon or an
active region is made of (the reference keeps these in core/actuators.py). Pixel arrays use
the layout of a body's own mask: (H, W), row 0 at the bottom, column index growing with x.

Everything here is vectorised numpy over pixels or edges. The order of the edge lists is part
of the contract (the pressure term sums over them, and the golden fixtures pin it): pixels in
row-major order of the mask, and round each pixel its four sides counter-clockwise starting
with the right-hand side.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

# The corners of a unit pixel counter-clockwise, starting at the bottom-right one. Side k runs
# from corner k to corner k+1 with the pixel on its left, so it faces the neighbour at the
# side's direction turned clockwise: right, top, left, bottom in that order.
_CORNER = np.array([[1, 0], [1, 1], [0, 1], [0, 0]], np.int64)
_SIDE_FROM = _CORNER
_SIDE_TO = np.roll(_CORNER, -1, axis=0)
_SIDE_FACES = np.stack([(_SIDE_TO - _SIDE_FROM)[:, 1], -(_SIDE_TO - _SIDE_FROM)[:, 0]], 1)


def enclosed_voids(mask):
    """Sealed cavities of a mask: void regions, connected through pixel sides (not corners),
    that do not reach the border of the grid. Returns (labels (H, W), [cavity labels]) with
    the labels of scipy.ndimage.label in increasing order (reference core/actuators.py:16-26)."""
    void = ~(np.asarray(mask) > 0)
    labels, n = ndimage.label(void, structure=ndimage.generate_binary_structure(2, 1))
    rim = np.zeros(void.shape, bool)
    rim[[0, -1], :] = True
    rim[:, [0, -1]] = True
    vented = np.unique(labels[rim])
    return labels, [int(k) for k in np.setdiff1d(np.arange(1, n + 1), vented)]


def _pixel_sides(pix_rc):
    """For pixels given as (row, col): the grid coordinates (ix, iy) of every side's start and
    end corner and of the neighbour it faces, shaped (n_pixels, 4, 2), sides in CCW order."""
    base = np.stack([pix_rc[:, 1], pix_rc[:, 0]], 1)[:, None, :]       # (ix, iy)
    return base + _SIDE_FROM[None], base + _SIDE_TO[None], base + _SIDE_FACES[None]


def _node_ids(node_of, corners):
    """Global node ids of grid corners (..., 2) through the body's corner -> node map."""
    flat = corners.reshape(-1, 2)
    return np.fromiter((node_of(int(ix), int(iy)) for ix, iy in flat), np.int64, len(flat))


def chamber_edges(mask, labels, label, node_of):
    """Directed wall edges of one cavity with the fluid on the left: every side a cavity pixel
    shares with a solid neighbour, oriented counter-clockwise round that pixel; pixels in
    row-major order, sides in the order of _CORNER (reference core/actuators.py:29-52)."""
    solid = np.asarray(mask) > 0
    H, W = solid.shape
    pix = np.argwhere(np.asarray(labels) == label)
    start, end, face = _pixel_sides(pix)
    fx, fy = face[..., 0], face[..., 1]
    inside = (fx >= 0) & (fx < W) & (fy >= 0) & (fy < H)
    wall = np.zeros(inside.shape, bool)
    wall[inside] = solid[fy[inside], fx[inside]]
    ea, eb = _node_ids(node_of, start[wall]), _node_ids(node_of, end[wall])
    if (ea < 0).any() or (eb < 0).any():
        raise RuntimeError("chamber wall node is not in the mesh")
    return ea, eb


def exterior_edges(mask, labels, enclosed, node_of):
    """Directed edges of a body's outer surface — solid pixel sides that face beyond the grid
    or a void that is not a sealed cavity — oriented with the surroundings on the left (each
    side reversed from its solid pixel's counter-clockwise order), so clockwise round the body
    (reference core/actuators.py:55-73)."""
    solid = np.asarray(mask) > 0
    labels = np.asarray(labels)
    H, W = solid.shape
    pix = np.argwhere(solid)
    start, end, face = _pixel_sides(pix)
    fx, fy = face[..., 0], face[..., 1]
    inside = (fx >= 0) & (fx < W) & (fy >= 0) & (fy < H)
    open_air = ~inside
    nb_lab = np.zeros(inside.shape, labels.dtype)
    nb_solid = np.zeros(inside.shape, bool)
    nb_lab[inside] = labels[fy[inside], fx[inside]]
    nb_solid[inside] = solid[fy[inside], fx[inside]]
    open_air |= inside & ~nb_solid & ~np.isin(nb_lab, np.asarray(list(enclosed), labels.dtype))
    return _node_ids(node_of, end[open_air]), _node_ids(node_of, start[open_air])


def loop_area(x, ea, eb, ref=None):
    """Signed area swept by the directed edges a -> b as seen from `ref` (default: the mean of
    the a ends): positive when they run counter-clockwise round it."""
    x = np.asarray(x, np.float64)
    o = x[ea].mean(0) if ref is None else np.asarray(ref, np.float64)
    pa, pb = x[ea] - o, x[eb] - o
    return float(0.5 * np.sum(pa[:, 0] * pb[:, 1] - pb[:, 0] * pa[:, 1]))


def mesh_boundary_loops(tri, x):
    """The boundary of a triangle mesh as closed loops of directed edges with the outside (the
    fluid) on the left: [(ea, eb, signed area)], holes counter-clockwise (> 0), the outline
    clockwise (< 0) (reference core/actuators.py:76-113).

    The half-edges of the counter-clockwise triangles are listed side 0 of every triangle,
    then side 1, then side 2; a half-edge whose edge no other triangle has is a boundary edge,
    and reversed it has the fluid on its left. Successor of a boundary edge: the one that
    starts where it ends. The loops are the cycles of that permutation, each listed from its
    first boundary edge, in the order of those first edges."""
    tri = np.asarray(tri, np.int64)
    x = np.asarray(x, np.float64)
    p = x[tri]
    twice = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
             - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    ccw = np.where((twice < 0)[:, None], tri[:, [0, 2, 1]], tri)
    half = np.concatenate([ccw[:, [k, (k + 1) % 3]] for k in range(3)])
    from .mesh import unique_pairs
    _, _, count, which = unique_pairs(half)
    bnd = half[count[which] == 1][:, ::-1]
    n = len(bnd)
    if n == 0:
        return []
    order = np.argsort(bnd[:, 0], kind="stable")
    starts = bnd[order, 0]
    clash = np.flatnonzero(starts[1:] == starts[:-1])
    if len(clash):
        raise ValueError(f"node {starts[clash[0]]} is where two boundary loops meet: the pressure "
                         f"boundary is ambiguous there. Separate them by an element, or join them.")
    pos = np.searchsorted(starts, bnd[:, 1])
    if (pos >= n).any() or (starts[np.minimum(pos, n - 1)] != bnd[:, 1]).any():
        raise ValueError("the mesh boundary does not close into loops")
    succ = order[pos]
    # cycle of every edge, named by its smallest edge index: pointer doubling
    root, jump = np.arange(n), succ.copy()
    for _ in range(int(np.ceil(np.log2(n))) + 1):
        root = np.minimum(root, root[jump])
        jump = jump[jump]
    # rank of every edge along its cycle, walking all cycles at once from their roots
    heads = np.unique(root)
    rank = np.zeros(n, np.int64)
    cur, live, k = heads.copy(), np.ones(len(heads), bool), 0
    while live.any():
        k += 1
        cur = succ[cur]
        live &= cur != heads
        rank[cur[live]] = k
    seq = np.lexsort((rank, root))
    cut = np.searchsorted(root[seq], heads)
    loops = []
    for ids in np.split(seq, cut[1:]):
        ea, eb = bnd[ids, 0], bnd[ids, 1]
        loops.append((ea, eb, loop_area(x, ea, eb)))
    return loops


def point_in_loop(point, x, ea, eb):
    """Is a point inside a closed loop of edges? Even-odd count of the edges a horizontal ray
    to the right of the point crosses."""
    px, py = float(point[0]), float(point[1])
    x = np.asarray(x, np.float64)
    a, b = x[ea], x[eb]
    spans = (a[:, 1] > py) != (b[:, 1] > py)
    a, b = a[spans], b[spans]
    t = (py - a[:, 1]) / (b[:, 1] - a[:, 1])
    return bool(np.count_nonzero(a[:, 0] + t * (b[:, 0] - a[:, 0]) > px) & 1)


def pixel_of(point, size, center, shape):
    """(row, col) of the mask pixel holding a world point, or None when it lies off the grid."""
    H, W = shape
    lo = (center[0] - 0.5 * size[0], center[1] - 0.5 * size[1])
    col = int(np.floor((point[0] - lo[0]) / (size[0] / W)))
    row = int(np.floor((point[1] - lo[1]) / (size[1] / H)))
    if 0 <= row < H and 0 <= col < W:
        return row, col
    return None


def snap_path(points, node_pos, node_ids):
    """The node nearest each waypoint, with a node repeated by consecutive waypoints kept once
    (reference core/actuators.py:143-149)."""
    P = np.atleast_2d(np.asarray(points, np.float64))
    X = np.asarray(node_pos, np.float64)
    d2 = (P[:, None, 0] - X[None, :, 0]) ** 2 + (P[:, None, 1] - X[None, :, 1]) ** 2
    ids = np.asarray(node_ids)[np.argmin(d2, axis=1)]
    fresh = np.ones(len(ids), bool)
    fresh[1:] = ids[1:] != ids[:-1]
    return ids[fresh]


def path_length(x, nodes):
    """Length of the polyline through the given nodes."""
    seg = np.diff(np.asarray(x, np.float64)[np.asarray(nodes)], axis=0)
    return float(np.sqrt((seg * seg).sum(1)).sum())


def _pixel_rc(tri_quad, tri_ids, ny):
    """(row, col) of the pixel each of a pixel body's elements came from: quad q = col*ny + row."""
    col, row = np.divmod(np.asarray(tri_quad)[np.asarray(tri_ids)], ny)
    return row, col


def region_elements(region, tri_quad, tri_ids, ny):
    """The elements (global ids, from tri_ids) of a pixel body whose pixel is set in `region`
    ((H, W))."""
    row, col = _pixel_rc(tri_quad, tri_ids, ny)
    return np.asarray(tri_ids)[(np.asarray(region) > 0)[row, col]]


def activation_matrix(stretch, transverse=1.0, fiber=None):
    """Rest-shape change A of an active region: isotropic scaling by `stretch`, or with a fiber
    (an angle in radians, or a 2-vector) `stretch` along it and `transverse` across it,
    A = R diag(stretch, transverse) R^T with R = [fiber | fiber turned a quarter turn]."""
    if fiber is None:
        return float(stretch) * np.eye(2)
    f = np.array([np.cos(fiber), np.sin(fiber)]) if np.isscalar(fiber) else np.asarray(fiber, np.float64)
    f = f / np.hypot(f[0], f[1])
    R = np.array([[f[0], -f[1]], [f[1], f[0]]])
    return R @ np.diag([float(stretch), float(transverse)]) @ R.T


def pixel_field_on_elements(field, tri_quad, tri_ids, ny):
    """Value of an (H, W) pixel field on each of a pixel body's elements."""
    row, col = _pixel_rc(tri_quad, tri_ids, ny)
    return np.asarray(field, np.float64)[row, col]
