"""The scene path's step loop as a HIP graph (diff_2d_fem_amd.graph.StepGraph) against the same
steps taken one simulate() at a time, on the grip scene (affine block, soft pads whose backs
are driven, IPC contact with friction): the jaws' targets change every step, so every replay
carries new staged targets.

* Two replays of 6 steps equal 12 eager steps: 1e-9 of the body size in float64, 1e-5 in
  mixed precision (the runs differ only in the order of float atomics).
* The same replay three times from one restored state, with eager work and synchronisations
  between the replays, gives the eager result every time.
* What a graph cannot hold is refused: the per-kernel path, and a replay after the step's
  parameters changed.
"""
import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator
from diff_2d_fem_amd.graph import StepGraph
from tests.scenes import close_jaws, grip

pytestmark = pytest.mark.gpu

KW = dict(max_iter=60, eps=0.0, conv_check=0)
DT = 1e-3
SIZE = 4e-3


def _models(precision):
    (ma, pa), (mb, pb) = grip(n_env=2), grip(n_env=2)
    ga = ImplicitIntegrator(path="scene", precision=precision, **KW)
    gb = ImplicitIntegrator(path="scene", precision=precision, **KW)
    for m, p, g in ((ma, pa, ga), (mb, pb, gb)):
        close_jaws(m, p, 1, n_close=8)
        g.simulate(m, DT)
    return (ma, pa, ga), (mb, pb, gb)


def _state(s):
    return {k: v.clone() for k, v in vars(s).items() if torch.is_tensor(v)}


def _restore(s, snap):
    for k, v in snap.items():
        getattr(s, k).copy_(v)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_graph_replays_equal_eager_steps(precision):
    bar = 1e-9 if precision == "double" else 1e-5
    (ma, pa, ga), (mb, pb, gb) = _models(precision)
    graph = StepGraph(gb, mb, DT, steps=6)
    for k in range(2, 14):
        close_jaws(ma, pa, k, n_close=8)
        ga.simulate(ma, DT)
    for first in (2, 8):
        graph.run(lambda k: close_jaws(mb, pb, k, n_close=8), first)
    torch.cuda.synchronize()
    err = np.abs(mb.get_x() - ma.get_x()).max() / SIZE
    print(f"{precision}: 12 steps in 2 replays vs eager: max |dx| / size = {err:.3e}", flush=True)
    assert err < bar
    assert abs(mb.time - ma.time) < 1e-12
    assert np.array_equal(gb.stats(mb)["iters"], ga.stats(ma)["iters"])


def test_replays_from_one_state_all_equal_the_eager_steps():
    (ma, pa, ga), (mb, pb, gb) = _models("double")
    graph = StepGraph(gb, mb, DT, steps=4)
    for k in range(2, 6):
        close_jaws(ma, pa, k, n_close=8)
        ga.simulate(ma, DT)
    x_ref = ma.get_x()
    snap, t0 = _state(mb._solver), mb.time
    for rep in range(3):
        _restore(mb._solver, snap)
        mb.time = t0
        junk = torch.randn(4096, device="cuda").sum()          # eager work between replays
        torch.cuda.synchronize()
        graph.run(lambda k: close_jaws(mb, pb, k, n_close=8), 2)
        torch.cuda.synchronize()
        err = np.abs(mb.get_x() - x_ref).max() / SIZE
        print(f"replay {rep}: max |dx| / size = {err:.3e} ({float(junk):.3f})", flush=True)
        assert err < 1e-9


def test_graph_refuses_what_it_cannot_hold():
    m, pads = grip(n_env=1)
    g = ImplicitIntegrator(path="plain", **KW)
    close_jaws(m, pads, 1)
    g.simulate(m, DT)
    with pytest.raises(ValueError, match="per-kernel path"):
        StepGraph(g, m, DT, steps=2)
    m2, pads2 = grip(n_env=1)
    g2 = ImplicitIntegrator(path="scene", **KW)
    close_jaws(m2, pads2, 1)
    g2.simulate(m2, DT)
    graph = StepGraph(g2, m2, DT, steps=2)
    graph.run(lambda k: close_jaws(m2, pads2, k), 2)
    m2._solver.iter_max += 1                                   # a parameter the graph captured
    with pytest.raises(RuntimeError, match="capture a new graph"):
        graph.replay()
