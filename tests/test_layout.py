"""The element kernel's node tiles (diff_2d_fem_amd.layout.element_tiles), on the CPU.

A numpy emulation of the two launches the tiles drive (the chunk pass: gather, per-chunk sums,
owned nodes added directly and halo sums written to partial slots; the halo pass: each halo
node adds its slots) gives the same per-node sums as adding every triangle's contributions
directly, and the layout's invariants hold: each triangle's local indices name its own nodes,
owned entries are exactly the nodes of one chunk, and every partial slot is summed once.
"""
import numpy as np
import pytest

from diff_2d_fem_amd.layout import ELEM_NODES_MAX, element_tiles
from diff_2d_fem_amd.workloads import _star_mesh


def _mesh(seed):
    rng = np.random.default_rng(seed)
    V, T, _ = _star_mesh(rng, 10e-3, 3000)
    perm = rng.permutation(len(T))               # scattered triangle order: larger halos
    return V, T[perm] if seed % 2 else T


@pytest.mark.parametrize("seed,chunk", [(0, 1024), (1, 512), (2, 256), (3, 64)])
def test_two_pass_sums_equal_direct_sums(seed, chunk):
    V, T = _mesh(seed)
    N = len(V)
    et = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in element_tiles(T, N, chunk=chunk).items()}
    off, nodes, hoff = et["node_off"], et["nodes"], et["halo_off"]
    nch = len(off) - 1
    assert et["max_nodes"] <= ELEM_NODES_MAX
    # local indices name the triangle's own nodes
    tl = et["tri_local"].astype(np.int64)
    loc = np.stack([tl & 1023, (tl >> 10) & 1023, (tl >> 20) & 1023], 1)
    cid = np.arange(len(T)) // chunk
    assert np.array_equal(nodes[off[cid][:, None] + loc], T)
    # owned entries: nodes of exactly one chunk, listed first in each chunk
    chunks_of = np.zeros(N, np.int64)
    for c in range(nch):
        chunks_of[np.unique(T[cid == c])] += 1
    for c in range(nch):
        lst = nodes[off[c]:off[c + 1]]
        no = len(lst) - (hoff[c + 1] - hoff[c])
        assert (chunks_of[lst[:no]] == 1).all() and (chunks_of[lst[no:]] > 1).all()
        assert (np.diff(lst[:no]) > 0).all() and (np.diff(lst[no:]) > 0).all()
    assert et["n_hnodes"] == int((chunks_of > 1).sum())
    # every slot lands in its own place of the node-sorted buffer
    hd = et["hdest"]
    assert np.array_equal(np.sort(hd), np.arange(hoff[-1]))
    # emulate both launches on a per-triangle vector contribution f (3 nodes x 2)
    rng = np.random.default_rng(seed + 10)
    f = rng.standard_normal((len(T), 3, 2))
    base = rng.standard_normal((N, 2))
    out = base.copy()
    part = np.zeros((max(hoff[-1], 1), 2))
    for c in range(nch):
        lst = nodes[off[c]:off[c + 1]]
        acc = np.zeros((len(lst), 2))
        ts = np.where(cid == c)[0]
        np.add.at(acc, loc[ts].reshape(-1), f[ts].reshape(-1, 2))
        no = len(lst) - (hoff[c + 1] - hoff[c])
        out[lst[:no]] += acc[:no]
        part[hd[hoff[c]:hoff[c + 1]]] = acc[no:]
    ho = et["hnode_off"]
    for k, n in enumerate(et["hnode"]):
        out[n] += part[ho[k]:ho[k + 1]].sum(0)
    want = base.copy()
    np.add.at(want, T.reshape(-1), f.reshape(-1, 2))
    assert np.abs(out - want).max() < 1e-12 * np.abs(want).max()


def test_lists_longer_than_the_limit_are_flagged():
    """Disjoint triangles: a 1024-triangle chunk touches 3072 nodes, beyond the 10-bit local
    indices; the tiles report it (max_nodes) so the launch takes the untiled kernel."""
    T = np.arange(3 * 2048).reshape(-1, 3)
    et = element_tiles(T, 3 * 2048)
    assert et["max_nodes"] == 3072 > ELEM_NODES_MAX


@pytest.mark.parametrize("chunk", [1024, 256])
def test_pressure_edges_dealt_to_tiles(chunk):
    """layout.pressure_tiles on the pneumatic finger's chamber edges: every edge lands in a chunk
    whose node list holds both its nodes at the local indices it records, the chunks' ranges
    partition the edges, and each edge keeps its chamber; an edge that is no triangle's side is
    refused."""
    from diff_2d_fem_amd import workloads as W
    from diff_2d_fem_amd.layout import pressure_tiles
    w = W.pneumatic_fingers(3)
    m = w["model"]
    me = m.me
    ea, eb, ech = [], [], []
    for i, c in enumerate(m.actuation.chambers):
        ea.append(np.asarray(c.ea, np.int64))
        eb.append(np.asarray(c.eb, np.int64))
        ech.append(np.full(len(c.ea), i))
    ea, eb, ech = np.concatenate(ea), np.concatenate(eb), np.concatenate(ech)
    N = me.rest_pos.shape[0]
    et = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in element_tiles(me.tri, N, chunk=chunk).items()}
    pt = pressure_tiles(me.tri, et["tri_local"], ea, eb, ech, N, chunk=chunk)
    off, loc, pch = pt["pe_off"], pt["pe_loc"].astype(np.int64), pt["pe_ch"]
    assert off[0] == 0 and off[-1] == len(ea) and (np.diff(off) >= 0).all()
    got = set()
    for c in range(len(off) - 1):
        lst = et["nodes"][et["node_off"][c]:et["node_off"][c + 1]]
        for q in range(off[c], off[c + 1]):
            a, b = lst[loc[q] & 1023], lst[(loc[q] >> 10) & 1023]
            got.add((int(a), int(b), int(pch[q])))
    assert got == set(zip(ea.tolist(), eb.tolist(), ech.tolist()))
    far = np.array([0]), np.array([N - 1]), np.array([0])
    assert pressure_tiles(me.tri, et["tri_local"], *far, N, chunk=chunk) is None
