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
ELEM_NODES_MAX = 2048      # longest node list one chunk sums in LDS (D2F_ELEM_NODES_MAX)


def element_tiles(tri, n_nodes, chunk=ELEM_CHUNK, device="cpu"):
    """Node tiles of the tiled element kernel. Triangle t belongs to chunk t // chunk. Returns
    torch tensors on `device`:
      node_off  (n_chunks + 1,) int32   chunk c's nodes are nodes[node_off[c]:node_off[c + 1]]
      nodes     int32                   ascending within a chunk
      owned     uint8                   1 when every triangle at that node is in that chunk
      tri_local (T, 4) int16 (read as uint16)  positions of each triangle's nodes in its
                                        chunk's list (column 3 is padding)
    and max_nodes, the longest list."""
    import torch
    t = torch.as_tensor(np.asarray(tri, np.int64).reshape(-1, 3), device=device)
    T = t.shape[0]
    nch = (T + chunk - 1) // chunk
    cid = torch.arange(T, device=device).div(chunk, rounding_mode="floor")
    key = (cid[:, None] * n_nodes + t).reshape(-1)
    uniq, inv = torch.unique(key, sorted=True, return_inverse=True)
    ucid = uniq.div(n_nodes, rounding_mode="floor")
    node = uniq - ucid * n_nodes
    off = torch.searchsorted(ucid, torch.arange(nch + 1, device=device))
    local = inv.reshape(T, 3) - off[cid][:, None]
    count = torch.bincount(node, minlength=n_nodes)
    owned = (count[node] == 1).to(torch.uint8)
    tl = torch.zeros((T, 4), dtype=torch.int32, device=device)
    tl[:, :3] = local.to(torch.int32)
    max_nodes = int((off[1:] - off[:-1]).max()) if nch else 0
    if max_nodes > 65535:
        raise ValueError("a chunk touches more than 65535 nodes")
    return dict(node_off=off.to(torch.int32), nodes=node.to(torch.int32), owned=owned,
                tri_local=tl.to(torch.int16), max_nodes=max_nodes)
