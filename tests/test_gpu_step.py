"""GPU parity of both device paths (per-kernel "plain" and scene-persistent "scene") against
the float64 CPU oracle.

Per-kernel entry points are called through the C ABI on identical inputs and must agree to
1e-12 relative (float64, different summation order only). Whole steps are run with a fixed
iteration count (eps = 0, no convergence poll) so that both sides take the same number of
iterations; positions must then agree to 1e-9 of the body size.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator, _lib
from diff_2d_fem_amd.layout import ELEM_NODES_MAX, element_tiles, tiles_struct
from oracle import terms as OT
from oracle.pncg import OracleIntegrator
from tests.scenes import strip, grip, close_jaws

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def deformed():
    """The strip mesh, randomly deformed (some elements compressed hard) with a random direction."""
    m = strip(n_env=3)
    me = m.me
    rng = np.random.default_rng(3)
    x = me.rest_pos + rng.normal(scale=1.5e-4, size=me.rest_pos.shape)
    p = rng.normal(scale=1e-4, size=me.rest_pos.shape)
    return me, x, p


def test_elastic_grad_diag(deformed):
    me, x, _ = deformed
    L = _lib.lib()
    d = torch.device("cuda")
    T = len(me.rest_area)
    xt = torch.tensor(x, device=d)
    tri = torch.tensor(me.tri, dtype=torch.int32, device=d)
    Bm = torch.tensor(me.D_m_inv.reshape(-1, 4), device=d)
    A, mu, lam = (torch.tensor(v, device=d) for v in (me.rest_area, me.mu, me.lam))
    et = element_tiles(me.tri, len(x), device=d)
    assert et["max_nodes"] > 1 and et["n_hnodes"] > 0 and et["owned"].bool().any()
    tiles = tiles_struct(et)
    # atomics only (no tiles) and tiled
    for tl in (None, tiles):
        for clamp in (1, 0):
            # the kernel accumulates: start from values at the scale of the elastic terms
            g0 = torch.tensor(1e-7 * np.random.default_rng(5).normal(size=x.shape), device=d)
            g, dg = g0.clone(), g0.clone()
            _lib.check(L.d2f_elastic_grad_diag_f64(_p(xt), _p(tri), _p(Bm), _p(A), _p(mu), _p(lam), T,
                                                   4e-6, clamp, None if tl is None else C.byref(tl),
                                                   _p(g), _p(dg), _stream()), "grad_diag")
            go, do = g0.cpu().numpy().copy(), g0.cpu().numpy().copy()
            OT.elastic_grad_diag(x, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 4e-6, clamp, go, do)
            assert _rel(g.cpu().numpy(), go) < 1e-12
            assert _rel(dg.cpu().numpy(), do) < 1e-12


def test_elastic_curvature_energy_caps_hvp(deformed):
    me, x, p = deformed
    L = _lib.lib()
    d = torch.device("cuda")
    T, B = len(me.rest_area), me.n_envs
    xt, pt = torch.tensor(x, device=d), torch.tensor(p, device=d)
    tri = torch.tensor(me.tri, dtype=torch.int32, device=d)
    t2e = torch.tensor(me.tri2env, dtype=torch.int32, device=d)
    Bm = torch.tensor(me.D_m_inv.reshape(-1, 4), device=d)
    A, mu, lam = (torch.tensor(v, device=d) for v in (me.rest_area, me.mu, me.lam))
    pHp = torch.zeros(B, dtype=torch.float64, device=d)
    _lib.check(L.d2f_elastic_pHp_f64(_p(xt), _p(pt), _p(tri), _p(Bm), _p(A), _p(mu), _p(lam), _p(t2e),
                                     T, 4e-6, 1, _p(pHp), _stream()), "pHp")
    po = np.zeros(B)
    OT.elastic_pHp(x, p, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 4e-6, 1, me.tri2env, po)
    assert _rel(pHp.cpu().numpy(), po) < 1e-12
    E = torch.zeros(B, dtype=torch.float64, device=d)
    _lib.check(L.d2f_elastic_energy_f64(_p(xt), _p(tri), _p(Bm), _p(A), _p(mu), _p(lam), _p(t2e), T,
                                        4e-6, _p(E), _stream()), "energy")
    Eo = np.zeros(B)
    OT.elastic_energy(x, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 4e-6, me.tri2env, Eo)
    assert _rel(E.cpu().numpy(), Eo) < 1e-12
    alpha = torch.full((B,), 3.0, dtype=torch.float64, device=d)
    pbig = pt * 40.0
    _lib.check(L.d2f_inversion_alpha_f64(_p(xt), _p(pbig), _p(tri), _p(t2e), T, 0.1, _p(alpha), _stream()),
               "inversion")
    ao = np.full(B, 3.0)
    OT.inversion_alpha(x, p * 40.0, me.tri, me.tri2env, 0.1, ao)
    assert (ao < 3.0).all()
    assert _rel(alpha.cpu().numpy(), ao) < 1e-12
    # the Hessian-vector product accumulates too: atomics only (no tiles), then tiled
    tiles = tiles_struct(element_tiles(me.tri, len(x), device=d))
    yo = np.zeros_like(x)
    OT.elastic_hvp(x, p, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 4e-6, yo)
    y0 = np.abs(yo).max() * np.random.default_rng(6).normal(size=x.shape)   # at the product's scale
    yo += y0
    for tl in (None, tiles):
        y = torch.tensor(y0, device=d)
        _lib.check(L.d2f_elastic_hvp_f64(_p(xt), _p(pt), _p(tri), _p(Bm), _p(A), _p(mu), _p(lam), T, 4e-6,
                                         None if tl is None else C.byref(tl), _p(y), _stream()), "hvp")
        assert _rel(y.cpu().numpy(), yo) < 1e-12


def _run_pair(make, n_steps, dt, path, kw, each=None, precision="double"):
    mo, mg = make(), make()
    if isinstance(mo, tuple):
        (mo, po), (mg, pg) = mo, mg
    else:
        po = pg = None
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path=path, precision=precision, **kw)
    for k in range(1, n_steps + 1):
        if each is not None:
            each(mo, po, k)
            each(mg, pg, k)
        oi.simulate(mo, dt)
        gi.simulate(mg, dt)
    return oi, gi, mo, mg


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_strip_steps_fixed_iterations(path):
    """Strip on the ground: gravity, ground barrier + friction, IPC self-contact, rigid modes."""
    kw = dict(max_iter=40, eps=0.0, conv_check=0)
    oi, gi, mo, mg = _run_pair(lambda: strip(n_env=2), 6, 2e-3, path, kw)
    xo, xg = oi.solver.x, mg.get_x()
    size = 0.054
    err = np.abs(xg - xo).max() / size
    print(f"strip {path}: max |dx| / size = {err:.3e}")
    assert err < 1e-9
    st = gi.stats(mg)
    assert (st["iters"] == oi.stats()["iters"]).all()


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_strip_steps_converged(path):
    """Default tolerances (eps = 1e-10, convergence poll every 8 iterations)."""
    oi, gi, mo, mg = _run_pair(lambda: strip(n_env=2), 8, 2e-3, path, {})
    err = np.abs(mg.get_x() - oi.solver.x).max() / 0.054
    print(f"strip converged {path}: max |dx| / size = {err:.3e}, iters {gi.stats(mg)['iters']} vs "
          f"{oi.stats()['iters']}")
    assert err < 1e-7


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_grip_steps_fixed_iterations(path):
    """Pads closing on a rigid block: IPC with friction between bodies, affine body, kinematic
    sweep, collider contact."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0)

    def each(m, pads, k):
        close_jaws(m, pads, k, n_close=8)
    oi, gi, mo, mg = _run_pair(lambda: grip(n_env=2), 12, 1e-3, path, kw, each)
    assert oi.solver.ipc.clr.min() < mo.dhat          # the pads are in contact with the block
    err = np.abs(mg.get_x() - oi.solver.x).max() / 8e-3
    print(f"grip {path}: max |dx| / size = {err:.3e}")
    assert err < 1e-9


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_grip_line_search(path):
    """The grip with the energy line search on: backtracking restarts the rigid block from its
    coordinates at the start of the iteration."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0, line_search=True, n_backtrack=4)

    def each(m, pads, k):
        close_jaws(m, pads, k, n_close=8)
    oi, gi, mo, mg = _run_pair(lambda: grip(n_env=2), 12, 1e-3, path, kw, each)
    err = np.abs(mg.get_x() - oi.solver.x).max() / 8e-3
    print(f"grip line search {path}: max |dx| / size = {err:.3e}")
    assert err < 1e-9


def test_scene_mixed_strip_close_to_float64():
    """Mixed precision (float32 element arithmetic) on the strip: positions within 1e-5 of the
    body size of the float64 oracle after the same fixed iterations."""
    kw = dict(max_iter=40, eps=0.0, conv_check=0)
    oi, gi, mo, mg = _run_pair(lambda: strip(n_env=2), 6, 2e-3, "scene", kw, precision="mixed")
    err = np.abs(mg.get_x() - oi.solver.x).max() / 0.054
    print(f"strip mixed: max |dx| / size = {err:.3e}")
    assert err < 1e-5


def test_scene_mixed_grip_close_to_float64():
    kw = dict(max_iter=60, eps=0.0, conv_check=0)

    def each(m, pads, k):
        close_jaws(m, pads, k, n_close=8)
    oi, gi, mo, mg = _run_pair(lambda: grip(n_env=2), 12, 1e-3, "scene", kw, each, precision="mixed")
    err = np.abs(mg.get_x() - oi.solver.x).max() / 8e-3
    print(f"grip mixed: max |dx| / size = {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("override", [dict(nacc_lds=0), dict(xc_lds=0), dict(excl_bits_lds=0),
                                      dict(max_pairs=0), dict(max_pairs=8),
                                      dict(nacc_lds=0, xc_lds=0, excl_bits_lds=0, max_pairs=0)])
def test_scene_fallback_paths(override, precision):
    """The scene kernel's fallbacks for scenes too large for LDS, forced on the grip scene:
    colour-by-colour element pass (nacc_lds=0), positions read from global memory
    (xc_lds=0), exclusion from the byte matrix (excl_bits_lds=0), grid traversal
    (max_pairs=0) and a pair list that overflows
    (max_pairs=8). Same bar as the default layout: 1e-9 of the body size in float64, 1e-5 in
    mixed precision."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0)

    def each(m, pads, k):
        close_jaws(m, pads, k, n_close=8)
    mo, po = grip(n_env=2)
    mg, pg = grip(n_env=2)
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path="scene", precision=precision, **kw)
    for k in range(1, 13):
        each(mo, po, k)
        each(mg, pg, k)
        oi.simulate(mo, 1e-3)
        if k == 1:
            gi.attach(mg, 1e-3)
            s = mg._solver
            s.lds_overrides = override
            s.build_layout()
            for key, v in override.items():
                assert getattr(s, key) <= v
        gi.simulate(mg, 1e-3)
    err = np.abs(mg.get_x() - oi.solver.x).max() / 8e-3
    print(f"{override} {precision}: max |dx| / size = {err:.3e}")
    assert err < (1e-9 if precision == "double" else 1e-5)


def test_scene_too_large_for_lds_takes_the_plain_path_alone():
    """A batch with one scene whose boundary overflows one workgroup's LDS (a 2 x 3000 px strip:
    about 6,000 boundary nodes) and two small ones: the large scene is stepped by the per-kernel
    path (float64) and the others keep the scene path and its precision. Each environment
    matches the same environment stepped by its own path in a batch of one kind (the order of
    float64 atomics aside)."""
    from diff_2d_fem_amd import Model

    def make(big=True):
        m = Model()
        m.add_ground(0.0, mu=0.5)
        for i in range(3):
            if i == 1 and big:
                m.add_soft_body(np.ones((2, 3000), np.uint8), physical_size=(1.5, 0.001),
                                center=(0.0, 0.0008), E=0.5e6, nu=0.40, density=1100.0, env_id=i)
            else:
                m.add_soft_body(np.ones((4, 12), np.uint8), physical_size=(0.012, 0.004),
                                center=(0.0, 0.0025), E=0.5e6, nu=0.40, density=1100.0, env_id=i)
        m.finalize()
        m.gravity = (0.0, -9.81)
        return m
    kw = dict(max_iter=20, eps=0.0, conv_check=0, self_collision=False)
    ms, mp, mr = make(), make(), make()
    gs = ImplicitIntegrator(path="scene", precision="double", **kw)
    gp = ImplicitIntegrator(path="plain", precision="double", **kw)
    with pytest.warns(UserWarning, match="per-kernel path"):
        gs.simulate(ms, 1e-3)
    gp.simulate(mp, 1e-3)
    s = ms._solver
    assert s.path == "scene" and list(s.plain_envs) == [1] and list(s.scene_envs) == [0, 2]
    assert list(gs.stats(ms)["plain_envs"]) == [1]
    for _ in range(2):
        gs.simulate(ms, 1e-3)
        gp.simulate(mp, 1e-3)
    env = np.asarray(ms.me.node2env)
    xs, xp = ms.get_x(), mp.get_x()
    assert np.abs(xs[env == 1] - xp[env == 1]).max() < 1e-12
    # the small scenes against a batch that holds only small scenes, all on the scene path
    small = make(big=False)
    gq = ImplicitIntegrator(path="scene", precision="double", **kw)
    for _ in range(3):
        gq.simulate(small, 1e-3)
    envq = np.asarray(small.me.node2env)
    for e in (0, 2):
        assert np.abs(xs[env == e] - small.get_x()[envq == e]).max() < 1e-12
    assert (gs.stats(ms)["iters"] == 20).all()


def _far_grip():
    """The grip scene with one more soft block 0.2 m away in the same environment: the
    environment's extent forces the plain path's contact grid to cells much wider than the
    reach."""
    from diff_2d_fem_amd import Model
    px, block = 0.5e-3, 8
    m = Model()
    m.body_friction = 0.8
    m.add_ground(0.0, mu=0.5)
    m.add_affine_body(np.ones((block, block), np.uint8), physical_size=(block * px, block * px),
                      center=(0.0, block * px / 2 + 0.9 * px / 8), density=600.0)
    pads = [m.add_soft_body(np.ones((8, 4), np.uint8), physical_size=(4 * px, 8 * px),
                            center=(sg * (block * px / 2 + 2 * px + 2 * px), 4 * px + 0.8e-3),
                            E=1.0e6, nu=0.40, density=1100.0) for sg in (-1, 1)]
    m.add_soft_body(np.ones((4, 4), np.uint8), physical_size=(4 * px, 4 * px), center=(0.2, 0.0021),
                    E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    m.gravity = (0.0, -9.81)
    for sg, p in zip((-1, 1), pads):
        X = m.rest_x(p)
        back = X[:, 0].min() if sg < 0 else X[:, 0].max()
        m.set_kinematic_constraint(p, np.where(np.abs(X[:, 0] - back) < 1e-9)[0])
    return m, [pads]


@pytest.mark.parametrize("scene", ["grip", "far"])
def test_ipc_pairs_equal_brute_force(scene):
    """The per-kernel path's grid broadphase finds exactly the oracle's brute-force pair set
    (every boundary node against every edge of its scene, same filters), at several states of
    a closing grip; "far" adds a body 0.2 m away, which coarsens the grid."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    m, pads = grip(n_env=2) if scene == "grip" else _far_grip()
    gi = ImplicitIntegrator(path="plain", precision="double", **kw)
    oi = OracleIntegrator(**kw)
    checked = 0
    for k in range(1, 13):
        close_jaws(m, pads, k, n_close=8)
        gi.simulate(m, 1e-3)
        if k % 3:
            continue
        s = m._solver
        got = s.ipc_pairs()
        if oi.solver is None:
            oi.attach(m, 1e-3)
        ipc = oi.solver.ipc
        K, JB = ipc.pairs(s.get_x())[:2]
        want = np.stack([K, ipc.bnd[JB]], 1)
        want = want[np.lexsort((want[:, 1], want[:, 0]))]
        assert len(want) > 0
        assert got.shape == want.shape and np.array_equal(got, want)
        checked += len(want)
    print(f"{scene}: {checked} pairs checked")


def test_elastic_chunks_beyond_the_tile_limit():
    """Chunks whose node lists exceed D2F_ELEM_NODES_MAX (512 single-pixel bodies: 2048 nodes
    per 1024-triangle chunk) take the untiled kernel, against the oracle."""
    from diff_2d_fem_amd.mesh import build_scene
    me = build_scene([dict(mask_xy=np.ones((1, 1), bool), size=(1e-3, 1e-3), center=(0.0, 0.0),
                           young=1e6, nu=0.4, rho=1000.0, env=i) for i in range(600)])
    rng = np.random.default_rng(4)
    x = me.rest_pos + rng.normal(scale=5e-5, size=me.rest_pos.shape)
    d = torch.device("cuda")
    et = element_tiles(me.tri, len(x), device=d)
    assert et["max_nodes"] > ELEM_NODES_MAX
    tiles = tiles_struct(et)
    xt = torch.tensor(x, device=d)
    tri = torch.tensor(me.tri, dtype=torch.int32, device=d)
    Bm = torch.tensor(me.D_m_inv.reshape(-1, 4), device=d)
    A, mu, lam = (torch.tensor(v, device=d) for v in (me.rest_area, me.mu, me.lam))
    g, dg = torch.zeros_like(xt), torch.zeros_like(xt)
    L = _lib.lib()
    _lib.check(L.d2f_elastic_grad_diag_f64(_p(xt), _p(tri), _p(Bm), _p(A), _p(mu), _p(lam), len(me.tri), 4e-6,
                                           1, C.byref(tiles), _p(g), _p(dg), _stream()), "grad_diag")
    go, do = np.zeros_like(x), np.zeros_like(x)
    OT.elastic_grad_diag(x, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 4e-6, 1, go, do)
    assert _rel(g.cpu().numpy(), go) < 1e-12 and _rel(dg.cpu().numpy(), do) < 1e-12


def test_elastic_grad_diag_with_halo_partials():
    """From 4,194,304 triangles on, the tiled element kernel writes its halo sums to partial
    slots and a second launch adds them per node (no global atomics): 4.3 M triangles of tiled
    star meshes against the oracle."""
    from diff_2d_fem_amd.workloads import _star_mesh
    Vs, Ts = [], []
    for i in range(6):
        V, Tm, _ = _star_mesh(np.random.default_rng([7, i]), 10e-3, 12000)
        Vs.append(V)
        Ts.append(Tm)
    reps = 4_300_000 // sum(len(t) for t in Ts) + 1
    nv = np.array([len(v) for v in Vs] * reps)
    off = np.concatenate([[0], np.cumsum(nv)[:-1]])
    tri = np.concatenate([Ts[k % 6] + off[k] for k in range(6 * reps)])
    X = np.concatenate([Vs[k % 6] for k in range(6 * reps)])
    T = len(tri)
    assert T >= 1 << 22
    D = np.stack([X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]]], 2)
    B = np.linalg.inv(D)
    area = 0.5 * np.abs(np.linalg.det(D))
    rng = np.random.default_rng(8)
    mu, lam = 1e5 * (1 + rng.random(T)), 4e5 * (1 + rng.random(T))
    x = X + rng.normal(scale=2e-5, size=X.shape)
    d = torch.device("cuda")
    et = element_tiles(tri, len(x), device=d)
    assert et["max_nodes"] <= ELEM_NODES_MAX and et["n_hnodes"] > 0
    tiles = tiles_struct(et)
    tt = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=d)
    g0 = rng.normal(scale=1e-7, size=x.shape)
    g, dg = tt(g0), tt(g0)
    args = [tt(x), tt(tri, torch.int32), tt(B.reshape(-1, 4)), tt(area), tt(mu), tt(lam)]   # kept alive
    L = _lib.lib()
    _lib.check(L.d2f_elastic_grad_diag_f64(*map(_p, args), T, 4e-6, 1, C.byref(tiles), _p(g), _p(dg), _stream()),
               "grad_diag")
    go, do = g0.copy(), g0.copy()
    OT.elastic_grad_diag(x, tri, B, area, mu, lam, 4e-6, 1, go, do)
    assert _rel(g.cpu().numpy(), go) < 1e-12 and _rel(dg.cpu().numpy(), do) < 1e-12
    # the Hessian-vector product through the same tiles and partial slots
    v = rng.normal(scale=1e-5, size=x.shape)
    yo = np.zeros_like(x)
    OT.elastic_hvp(x, v, tri, B, area, mu, lam, 4e-6, yo)
    y0 = np.abs(yo).max() * rng.normal(size=x.shape)        # at the product's scale
    yo += y0
    y = tt(y0)
    args.append(tt(v))
    _lib.check(L.d2f_elastic_hvp_f64(*map(_p, args[:1]), _p(args[-1]), *map(_p, args[1:6]), T, 4e-6, C.byref(tiles),
                                     _p(y), _stream()), "hvp")
    assert _rel(y.cpu().numpy(), yo) < 1e-12
