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
ELEM_NODES_MAX = 1024      # longest node list one chunk sums in LDS (D2F_ELEM_NODES_MAX; 10-bit indices)


def element_tiles(tri, n_nodes, chunk=ELEM_CHUNK, device="cpu"):
    """Node tiles of the tiled element kernel. Triangle t belongs to chunk t // chunk. A node
    is owned by a chunk when every triangle at it is in that chunk; the others are halo nodes,
    and every chunk touching one sums its part into a slot of a partial buffer that a second
    pass adds up per node, in chunk order (no global atomics). Returns torch tensors on `device`:
      node_off  (n_chunks + 1,) int32   chunk c's list is nodes[node_off[c]:node_off[c + 1]]
      nodes     int32                   per chunk its owned nodes, then its halo nodes, each
                                        part ascending
      halo_off  (n_chunks + 1,) int32   chunk c's halo entries (the last halo_off[c + 1] -
                                        halo_off[c] of its list) are partial slots
                                        halo_off[c] .. halo_off[c + 1]
      tri_local (T,) int32              positions of triangle t's nodes in its chunk's list,
                                        10 bits each (bits 0-9, 10-19, 20-29)
      hnode, hnode_off int32            halo node k = hnode[k] sums the partials
                                        part[hnode_off[k]:hnode_off[k + 1]] (chunk order)
      hdest     int32                   chunk-order slot s of a halo entry -> its position in
                                        that node-sorted partial buffer
      part      float64 (4 n_slots,)    the partial buffer (grad x, grad y, diag x, diag y)
      owned     uint8                   1 per owned list entry (diagnostics)
    and max_nodes (the longest list), n_hnodes. When a list exceeds ELEM_NODES_MAX nodes the
    kernel is not tiled (tri_local is then left zero)."""
    import torch
    t = torch.as_tensor(np.asarray(tri, np.int64).reshape(-1, 3), device=device)
    T = t.shape[0]
    nch = (T + chunk - 1) // chunk
    cid = torch.arange(T, device=device).div(chunk, rounding_mode="floor")
    key = (cid[:, None] * n_nodes + t).reshape(-1)
    uniq, inv = torch.unique(key, sorted=True, return_inverse=True)
    ucid = uniq.div(n_nodes, rounding_mode="floor")
    node = uniq - ucid * n_nodes
    count = torch.bincount(node, minlength=n_nodes)
    halo = count[node] > 1
    # owned entries first, halo entries after, within each chunk (stable: ascending node ids)
    perm = torch.argsort(ucid * 2 + halo.to(torch.int64), stable=True)
    pos = torch.empty_like(perm)
    pos[perm] = torch.arange(perm.numel(), device=device)
    node, ucid, halo = node[perm], ucid[perm], halo[perm]
    ar = torch.arange(nch + 1, device=device)
    off = torch.searchsorted(ucid, ar)
    local = pos[inv].reshape(T, 3) - off[cid][:, None]
    max_nodes = int((off[1:] - off[:-1]).max()) if nch else 0
    if max_nodes <= ELEM_NODES_MAX:
        packed = local[:, 0] | (local[:, 1] << 10) | (local[:, 2] << 20)
    else:
        packed = torch.zeros(T, dtype=torch.int64, device=device)
    hcount = torch.bincount(ucid[halo], minlength=nch)
    halo_off = torch.zeros(nch + 1, dtype=torch.int64, device=device)
    halo_off[1:] = torch.cumsum(hcount, 0)
    hnodes_slot = node[halo]                          # node of partial slot s (slots in chunk order)
    order = torch.argsort(hnodes_slot, stable=True)   # per node, its slots in chunk order
    hnode, hcnt = torch.unique_consecutive(hnodes_slot[order], return_counts=True)
    hnode_off = torch.zeros(hnode.numel() + 1, dtype=torch.int64, device=device)
    hnode_off[1:] = torch.cumsum(hcnt, 0)
    hdest = torch.empty_like(order)
    hdest[order] = torch.arange(order.numel(), device=device)
    i32 = torch.int32
    return dict(node_off=off.to(i32), nodes=node.to(i32), halo_off=halo_off.to(i32),
                tri_local=packed.to(i32), hnode=hnode.to(i32), hnode_off=hnode_off.to(i32),
                hdest=hdest.to(i32), part=torch.zeros(4 * max(int(halo_off[-1]), 1), dtype=torch.float64,
                                                      device=device),
                owned=(~halo).to(torch.uint8), max_nodes=max_nodes, n_hnodes=int(hnode.numel()))


def pressure_tiles(tri, tri_local, pe_a, pe_b, pe_ch, n_nodes, chunk=ELEM_CHUNK):
    """The pressure chamber edges dealt to the element tiles (d2f_elastic_pressure_hvp_f64):
    each edge goes to the chunk of a triangle having it as a side, with its two nodes as entries
    of that chunk's node list. Returns numpy arrays pe_off (n_chunks + 1,) int32, pe_loc uint32
    (lo | hi << 10) and pe_ch int32, edges grouped by chunk; None when an edge is no triangle's
    side."""
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    T = len(tri)
    a, b = np.asarray(pe_a, np.int64), np.asarray(pe_b, np.int64)
    s0, s1 = tri, tri[:, [1, 2, 0]]
    key = (np.minimum(s0, s1) * n_nodes + np.maximum(s0, s1)).reshape(-1)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    q = np.minimum(a, b) * n_nodes + np.maximum(a, b)
    pos = np.searchsorted(ks, q)
    if len(ks) == 0 or (pos >= len(ks)).any() or (ks[np.minimum(pos, len(ks) - 1)] != q).any():
        return None
    t = order[pos] // 3
    loc = np.asarray(tri_local, np.int64)[t]
    slots = np.stack([loc & 1023, (loc >> 10) & 1023, (loc >> 20) & 1023], 1)
    rows = np.arange(len(t))
    la = slots[rows, (tri[t] == a[:, None]).argmax(1)]
    lb = slots[rows, (tri[t] == b[:, None]).argmax(1)]
    c = t // chunk
    o = np.argsort(c, kind="stable")
    pe_off = np.searchsorted(c[o], np.arange((T + chunk - 1) // chunk + 1)).astype(np.int32)
    return dict(pe_off=pe_off, pe_loc=(la | (lb << 10))[o].astype(np.uint32),
                pe_ch=np.asarray(pe_ch, np.int64)[o].astype(np.int32))


def tiles_struct(et):
    """The d2f_elem_tiles struct of element_tiles' tensors."""
    from ._lib import ElemTiles
    st = ElemTiles()
    for k in ("node_off", "nodes", "halo_off", "tri_local", "hnode", "hnode_off", "hdest", "part"):
        setattr(st, k, et[k].data_ptr())
    st.max_nodes, st.n_hnodes = et["max_nodes"], et["n_hnodes"]
    return st
