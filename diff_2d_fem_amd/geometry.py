"""Host geometry behind the actuators: which mesh entities a pressure chamber, a tendon or
a region is made of (reference core/actuators.py). Pixel arrays use the layout of the body's
own mask: (H, W), row 0 at the bottom.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

_FOUR = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], bool)
# (neighbour offset) -> (edge start, edge end) as grid-corner offsets, counter-clockwise
# round a pixel: right wall going up, top going left, left going down, bottom going right
_WALLS = (((1, 0), ((1, 0), (1, 1))), ((0, 1), ((1, 1), (0, 1))),
          ((-1, 0), ((0, 1), (0, 0))), ((0, -1), ((0, 0), (1, 0))))


def enclosed_voids(mask):
    """Sealed cavities (4-connected void regions not reaching the grid edge): (labels, ids)
    (reference core/actuators.py:16-26)."""
    solid = np.asarray(mask) > 0
    lab, n = ndimage.label(~solid, structure=_FOUR)
    border = set(np.unique(np.concatenate([lab[0], lab[-1], lab[:, 0], lab[:, -1]]))) - {0}
    return lab, [k for k in range(1, n + 1) if k not in border]


def chamber_edges(mask, labels, label, node_of):
    """Directed wall edges of one cavity with the fluid on the left, pixel by pixel in
    row-major order of the cavity's pixels (reference core/actuators.py:29-52)."""
    solid = np.asarray(mask) > 0
    H, W = solid.shape
    ea, eb = [], []
    for iy, ix in zip(*np.nonzero(labels == label)):
        for (dx, dy), (a, b) in _WALLS:
            jx, jy = ix + dx, iy + dy
            if 0 <= jx < W and 0 <= jy < H and solid[jy, jx]:
                na, nb = node_of(ix + a[0], iy + a[1]), node_of(ix + b[0], iy + b[1])
                if na < 0 or nb < 0:
                    raise RuntimeError("chamber wall node is not in the mesh")
                ea.append(na)
                eb.append(nb)
    return np.asarray(ea, np.int64), np.asarray(eb, np.int64)


def exterior_edges(mask, labels, enclosed, node_of):
    """Directed edges of the outer surface, surroundings on the left (clockwise round the
    body) (reference core/actuators.py:55-73)."""
    solid = np.asarray(mask) > 0
    H, W = solid.shape
    enclosed = set(enclosed)
    ea, eb = [], []
    for iy, ix in zip(*np.nonzero(solid)):
        for (dx, dy), (a, b) in _WALLS:
            jx, jy = ix + dx, iy + dy
            outside = not (0 <= jx < W and 0 <= jy < H)
            if outside or (not solid[jy, jx] and labels[jy, jx] not in enclosed):
                ea.append(node_of(ix + b[0], iy + b[1]))
                eb.append(node_of(ix + a[0], iy + a[1]))
    return np.asarray(ea, np.int64), np.asarray(eb, np.int64)


def loop_area(x, ea, eb, ref=None):
    """Signed area enclosed by directed edges a -> b, about `ref` (default: mean of the a's)."""
    x = np.asarray(x, np.float64)
    c = x[ea].mean(0) if ref is None else np.asarray(ref, np.float64)
    ra, rb = x[ea] - c, x[eb] - c
    return float(0.5 * (ra[:, 0] * rb[:, 1] - ra[:, 1] * rb[:, 0]).sum())


def mesh_boundary_loops(tri, x):
    """Boundary loops of a triangle mesh as directed edges, fluid on the left:
    [(ea, eb, signed area)], holes > 0 (reference core/actuators.py:76-113)."""
    tri = np.asarray(tri, np.int64)
    x = np.asarray(x, np.float64)
    p = x[tri]
    cw = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
          - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])) < 0
    t = tri.copy()
    t[cw, 1], t[cw, 2] = tri[cw, 2], tri[cw, 1]
    d = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    _, inv, cnt = np.unique(np.sort(d, 1), axis=0, return_inverse=True, return_counts=True)
    fl = d[cnt[inv.reshape(-1)] == 1][:, ::-1]
    starts, n_out = np.unique(fl[:, 0], return_counts=True)
    if (n_out > 1).any():
        raise ValueError(f"node {starts[np.argmax(n_out > 1)]} is where two boundary loops meet: "
                         f"the pressure boundary is ambiguous there")
    nxt = dict(zip(fl[:, 0].tolist(), range(len(fl))))
    seen = np.zeros(len(fl), bool)
    loops = []
    for e0 in range(len(fl)):
        if seen[e0]:
            continue
        ids, e = [], e0
        while not seen[e]:
            seen[e] = True
            ids.append(e)
            e = nxt[int(fl[e, 1])]
        ea, eb = fl[ids, 0], fl[ids, 1]
        loops.append((ea, eb, loop_area(x, ea, eb)))
    return loops


def point_in_loop(point, x, ea, eb):
    """Even-odd test of a point against a closed loop of edges."""
    px, py = float(point[0]), float(point[1])
    a, b = np.asarray(x, np.float64)[ea], np.asarray(x, np.float64)[eb]
    up = (a[:, 1] > py) != (b[:, 1] > py)
    xc = a[up, 0] + (py - a[up, 1]) * (b[up, 0] - a[up, 0]) / (b[up, 1] - a[up, 1])
    return bool((xc > px).sum() % 2)


def pixel_of(point, size, center, shape):
    """(row, col) of the pixel holding a world point, or None off the grid."""
    H, W = shape
    col = int(np.floor((point[0] - (center[0] - 0.5 * size[0])) / (size[0] / W)))
    row = int(np.floor((point[1] - (center[1] - 0.5 * size[1])) / (size[1] / H)))
    return (row, col) if (0 <= row < H and 0 <= col < W) else None


def snap_path(points, node_pos, node_ids):
    """Nearest node to each waypoint, consecutive repeats dropped
    (reference core/actuators.py:143-149)."""
    points = np.atleast_2d(np.asarray(points, np.float64))
    d = np.linalg.norm(np.asarray(node_pos)[None, :, :] - points[:, None, :], axis=2)
    ids = np.asarray(node_ids)[d.argmin(1)]
    keep = np.concatenate([[True], ids[1:] != ids[:-1]])
    return ids[keep]


def path_length(x, nodes):
    p = np.asarray(x, np.float64)[np.asarray(nodes)]
    return float(np.linalg.norm(np.diff(p, axis=0), axis=1).sum())


def region_elements(region, tri_quad, tri_ids, ny):
    """Global element ids of a pixel body whose pixel lies in `region` ((H, W) bool)."""
    region = np.asarray(region) > 0
    q = np.asarray(tri_quad)[tri_ids]
    return np.asarray(tri_ids)[region[q % ny, q // ny]]


def activation_matrix(stretch, transverse=1.0, fiber=None):
    """Rest-shape change of an active region: isotropic, or `stretch` along a fiber and
    `transverse` across it."""
    if fiber is None:
        return np.eye(2) * float(stretch)
    f = np.asarray([np.cos(fiber), np.sin(fiber)] if np.isscalar(fiber) else fiber, np.float64)
    f = f / np.linalg.norm(f)
    g = np.array([-f[1], f[0]])
    return float(stretch) * np.outer(f, f) + float(transverse) * np.outer(g, g)


def pixel_field_on_elements(field, tri_quad, tri_ids, ny):
    """Value of an (H, W) pixel field on each of a body's elements."""
    q = np.asarray(tri_quad)[tri_ids]
    return np.asarray(field, np.float64)[q % ny, q // ny]
