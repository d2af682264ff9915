"""The host scene layer reproduces the reference's mesh indexing bit for bit (golden fixtures
written from the reference's own core/mesh.py and core/actuators.py by tests/golden/make_fixtures.py)."""
import numpy as np
import pytest

from diff_2d_fem_amd import Model, geometry as G
from diff_2d_fem_amd.mesh import build_scene, check_mesh, mask_to_mesh
from diff_2d_fem_amd import layout as LY
from tests.golden.fixture_scenes import scenes, MESH_KEYS

FX = np.load("tests/golden/mesh_fixtures.npz")
AX = np.load("tests/golden/actuator_fixtures.npz")


@pytest.mark.parametrize("name", sorted(scenes()))
def test_scene_arrays_are_bit_exact(name):
    me = build_scene([dict(b) for b in scenes()[name]])
    for k in MESH_KEYS:
        got = np.asarray(getattr(me, k))
        want = FX[f"{name}/{k}"]
        assert got.shape == want.shape and np.array_equal(got, want), f"{name}: {k} differs"


def test_strip_sizes():
    me = build_scene([dict(b) for b in scenes()["strip"]])
    assert me.rest_pos.shape[0] == 511 and len(me.tri) == 864 and len(me.boundary_nodes) == 156
    assert len(me.boundary_edges) == 156


def test_chamber_and_exterior_edges_and_tendon_snapping():
    ring = scenes()["multi"][0]["mask_xy"].T
    lab, ids = G.enclosed_voids(ring)
    assert np.array_equal(lab, AX["labels"]) and list(ids) == list(AX["ids"])
    me = build_scene([dict(scenes()["multi"][0], env=0)])
    remap = me.grid_to_live_node_per_body[0]
    ny = ring.shape[0]

    def node_of(ix, iy):
        loc = int(remap[(ny + 1) * ix + iy])
        return -1 if loc < 0 else loc
    ea, eb = G.chamber_edges(ring, lab, ids[0], node_of)
    assert np.array_equal(ea, AX["chamber_a"]) and np.array_equal(eb, AX["chamber_b"])
    xa, xb = G.exterior_edges(ring, lab, ids, node_of)
    assert np.array_equal(xa, AX["exterior_a"]) and np.array_equal(xb, AX["exterior_b"])
    path = G.snap_path(AX["tendon_points"], me.rest_pos, np.arange(me.rest_pos.shape[0]))
    assert np.array_equal(path, AX["tendon_path"])


def test_mask_to_mesh_matches_pixel_body():
    mask = np.ones((3, 5), np.uint8)
    mask[1, 2] = 0
    V, Tr, px = mask_to_mesh(mask, (5e-3, 3e-3), (0.0, 0.0))
    m = Model()
    h = m.add_soft_body(mask, physical_size=(5e-3, 3e-3), center=(0.0, 0.0))
    m.finalize()
    assert np.array_equal(V, m.rest_x(h)) and np.array_equal(Tr, m.me.tri)
    assert (mask[px[:, 0], px[:, 1]] > 0).all()


def test_check_mesh_refuses_defects():
    V = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], float)
    T = np.array([[0, 1, 2], [1, 3, 2]])
    info = check_mesh(V, T)
    assert info["L_max"] > 0
    with pytest.raises(ValueError):
        check_mesh(V, np.array([[0, 1, 1]]))
    with pytest.raises(ValueError):
        check_mesh(V, np.array([[0, 1, 2], [0, 1, 2]]))
    with pytest.raises(ValueError):
        check_mesh(np.vstack([V, [[5, 5]]]), T)


def test_overlapping_bodies_are_refused():
    m = Model()
    for _ in range(2):
        m.add_soft_body(np.ones((8, 8), np.uint8), physical_size=(0.008, 0.008), center=(0.0, 0.0))
    with pytest.raises(ValueError, match="env_id"):
        m.finalize()


@pytest.mark.parametrize("name", ["strip", "blob", "multi", "meshbody"])
def test_coloring_is_valid(name):
    me = build_scene([dict(b) for b in scenes()[name]])
    c = LY.color_mesh(me.tri, me.rest_pos.shape[0])
    assert LY.check_coloring(me.tri, c) and c.max() < LY.MAX_COLORS
    if name == "strip":
        cp = LY.color_pixel_body(me.tri_quad, 6)
        assert LY.check_coloring(me.tri, cp) and cp.max() == 7


def test_model_refuses_cpu_device():
    with pytest.raises(ValueError, match="GPU"):
        Model(device="cpu")


def test_element_tiles_layout():
    """Node tiles of the element kernel on a multi-body scene, checked by brute force: each
    chunk's list is exactly the set of nodes its triangles touch (owned ones first, then the
    halo, each ascending), a node is owned by a chunk iff no other chunk touches it, and the
    local indices point back at the triangle's nodes."""
    from diff_2d_fem_amd.layout import element_tiles
    me = build_scene([dict(b) for b in scenes()["multi"]])
    tri = np.asarray(me.tri)
    N = me.rest_pos.shape[0]
    ch = 37
    et = element_tiles(tri, N, chunk=ch)
    off, nodes = et["node_off"].numpy(), et["nodes"].numpy()
    owned, hoff = et["owned"].numpy(), et["halo_off"].numpy()
    tl = et["tri_local"].numpy().astype(np.int64)
    loc = np.stack([tl & 1023, (tl >> 10) & 1023, (tl >> 20) & 1023], 1)
    nch = (len(tri) + ch - 1) // ch
    assert len(off) == nch + 1
    chunks_at = [set() for _ in range(N)]
    for t, tr in enumerate(tri):
        for n in tr:
            chunks_at[n].add(t // ch)
    for c in range(nch):
        sub = tri[c * ch:(c + 1) * ch]
        lst = nodes[off[c]:off[c + 1]]
        assert np.array_equal(np.sort(lst), np.unique(sub))
        no = len(lst) - (hoff[c + 1] - hoff[c])
        for k, n in enumerate(lst):
            assert owned[off[c] + k] == (chunks_at[n] == {c}) == (k < no)
        assert np.array_equal(lst[loc[c * ch:(c + 1) * ch]], sub)
    assert et["max_nodes"] == int(np.diff(off).max())


def test_reference_table_scenes_match_the_published_sizes():
    """The throughput-table scenes (workloads.TABLE_SCENES, after the reference's
    tools/benchmark_envs.py) give the node counts and iteration budgets the reference's README
    quotes for them (README.md:206-208: beam 729 nodes and 177 iterations per step, strip 511
    and 167, the gripped block 595 nodes)."""
    from diff_2d_fem_amd import ImplicitIntegrator
    from diff_2d_fem_amd.workloads import TABLE_SCENES
    want = {"beam": (729, 177), "strip": (511, 167), "grip": (595, 177)}
    for name, build in TABLE_SCENES.items():
        m, kw, command, dt, lead = build(2)
        command(1)
        command(lead)
        assert m.n_nodes == 2 * want[name][0], name
        assert ImplicitIntegrator(**kw)._iters_for(m, dt) == want[name][1], name
