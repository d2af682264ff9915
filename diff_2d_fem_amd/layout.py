"""Host-side layout of the scene-persistent path: triangle colouring and per-scene offsets.

A colour class is a set of triangles no two of which share a node, so one workgroup can add
every element's contribution to its nodes with plain read-modify-write, colour after colour,
with no atomics and in a fixed order.
"""
from __future__ import annotations

import numpy as np

MAX_COLORS = 24


def color_pixel_body(tri_quad, ny):
    """Colours of a pixel body's triangles: quad parity (ix mod 2, iy mod 2) times the two
    triangles of a quad — 8 colours. Quads two apart in x or y share no grid corner."""
    q = np.asarray(tri_quad, np.int64)
    ix, iy = q // ny, q % ny
    second = np.zeros(len(q), np.int64)
    second[1:] = (q[1:] == q[:-1]).astype(np.int64)          # the quad's second triangle
    return ((ix % 2) * 2 + (iy % 2)) * 2 + second


def color_mesh(tri, n_nodes, seed=0):
    """Colours of a general triangle mesh (Jones-Plassmann with first fit): each round takes
    the uncoloured triangles whose priority is the largest among the uncoloured triangles at
    each of their three nodes — an independent set — and gives each the smallest colour none
    of its coloured neighbours has."""
    tri = np.asarray(tri, np.int64)
    T = len(tri)
    rng = np.random.default_rng(seed)
    prio = rng.permutation(T).astype(np.int64) + 1
    color = np.full(T, -1, np.int64)
    used = np.zeros(n_nodes, np.int64)                  # bit c: a triangle of colour c touches the node
    flat = tri.reshape(-1)
    while (color < 0).any():
        live = color < 0
        pr = np.where(live, prio, 0)
        node_max = np.zeros(n_nodes, np.int64)
        np.maximum.at(node_max, flat, np.repeat(pr, 3))
        win = np.nonzero(live & (pr[:, None] == node_max[tri]).all(1))[0]
        forbid = used[tri[win, 0]] | used[tri[win, 1]] | used[tri[win, 2]]
        c = np.zeros(len(win), np.int64)
        free = (forbid & 1) != 0
        while free.any():
            c[free] += 1
            free = ((forbid >> c) & 1) != 0
        if c.max(initial=0) >= MAX_COLORS:
            raise RuntimeError(f"triangle colouring needs more than {MAX_COLORS} colours")
        color[win] = c
        np.bitwise_or.at(used, tri[win].reshape(-1), np.repeat(np.int64(1) << c, 3))
    return color


def check_coloring(tri, color):
    """True when no two triangles of one colour share a node."""
    tri = np.asarray(tri, np.int64)
    for c in np.unique(color):
        nodes = tri[color == c].reshape(-1)
        if len(np.unique(nodes)) != len(nodes):
            return False
    return True


def starts_of(keys, n):
    """CSR offsets of items grouped by key in 0..n-1 (keys must be sorted)."""
    return np.searchsorted(np.asarray(keys), np.arange(n + 1), side="left").astype(np.int32)


ELEM_CHUNK = 1024          # triangles per workgroup of the tiled element kernel (D2F_ELEM_CHUNK)
ELEM_WINDOW_MAX = 4096     # largest node window a chunk accumulates in LDS


def element_owners(tri, n_nodes, chunk=ELEM_CHUNK):
    """Node ownership for the tiled element kernel. Triangle t belongs to chunk t // chunk;
    owner[n] is that chunk when every triangle at node n belongs to it (the chunk may then
    add into the node with a plain read-modify-write), else -1 (atomics). Also returns each
    chunk's node range (smallest and largest node index of its triangles, interleaved) and
    the window: the largest range size, capped at ELEM_WINDOW_MAX (a chunk with a wider
    range falls back to atomics)."""
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    T = len(tri)
    if T == 0:
        return np.full(n_nodes, -1, np.int32), np.zeros(0, np.int32), 1
    c = np.repeat(np.arange(T, dtype=np.int64) // chunk, 3)
    flat = tri.reshape(-1)
    lo = np.full(n_nodes, np.iinfo(np.int64).max)
    hi = np.full(n_nodes, -1, np.int64)
    np.minimum.at(lo, flat, c)
    np.maximum.at(hi, flat, c)
    owner = np.where(lo == hi, hi, -1).astype(np.int32)
    starts = np.arange(0, T, chunk)
    wmin = np.minimum.reduceat(tri.min(1), starts)
    wmax = np.maximum.reduceat(tri.max(1), starts)
    rng = np.stack([wmin, wmax], 1).reshape(-1).astype(np.int32)
    window = int(min((wmax - wmin + 1).max(), ELEM_WINDOW_MAX))
    return owner, rng, window
