"""The scene kernel's autotuner (diff_2d_fem_amd.tune) on the grip scene (affine block, soft pads
whose backs are driven, IPC contact with friction).

* Every plan on the tuner's axes steps to the same state as the default plan (1e-9 of the body
  size in float64): the grid and the pair list are exact broadphases, the workgroup size
  changes only reduction order.
* ImplicitIntegrator(autotune=True) measures candidates, applies the fastest, writes the cache;
  a second model of the same workload takes the plan from the cache without measuring.
"""
import json

import numpy as np
import pytest

from diff_2d_fem_amd import ImplicitIntegrator
from tests.scenes import close_jaws, grip

pytestmark = pytest.mark.gpu

KW = dict(max_iter=60, eps=0.0, conv_check=0)
DT = 1e-3
SIZE = 4e-3


def _run(plan=None, steps=10, **kw):
    m, pads = grip(n_env=4)
    g = ImplicitIntegrator(path="scene", **KW, **kw)
    for k in range(1, steps + 1):
        close_jaws(m, pads, k, n_close=8)
        if k == 1 and plan is not None:
            g.attach(m, DT)
            m._solver.build_layout()
            m._solver.set_plan(**plan)
        g.simulate(m, DT)
    return m, g


@pytest.mark.parametrize("plan", [dict(block_size=256), dict(block_size=1024), dict(max_cells=256),
                                  dict(max_cells=2048), dict(skin=0.25), dict(skin=1.0),
                                  dict(lds_order=("bits", "pairs", "nacc", "xc")),
                                  dict(block_size=1024, max_cells=256, skin=1.0)])
def test_every_plan_steps_to_the_same_state(plan):
    ref, _ = _run()
    m, _ = _run(plan)
    err = np.abs(m.get_x() - ref.get_x()).max() / SIZE
    print(f"{plan}: {m._solver.plan()}: max |dx| / size = {err:.3e}", flush=True)
    assert err < 1e-9


def test_autotune_applies_and_caches_a_plan(tmp_path, monkeypatch):
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("D2FEM_TUNE_CACHE", str(cache))
    ref, _ = _run()
    m, g = _run(autotune=True)
    t = g.tuned
    print(f"tuned: {t['block_size']} threads, {t['max_cells']} cells, skin {t['skin']}, "
          f"{t['lds_order']}; {len(t['trials'])} trials: "
          + ", ".join(f"{x['ms']:.3f}" for x in t["trials"]), flush=True)
    assert not t["cached"] and len(t["trials"]) >= 6
    assert round(t["ms_per_step"], 4) == min(x["ms"] for x in t["trials"])
    err = np.abs(m.get_x() - ref.get_x()).max() / SIZE
    assert err < 1e-9
    table = json.loads(cache.read_text())
    assert t["key"] in table and table[t["key"]]["shape"]["B"] == 4
    # the same workload again: the plan comes from the cache
    m2, g2 = _run(steps=1, autotune=True)
    assert g2.tuned["cached"] and g2.tuned["key"] == t["key"]
    for k in ("block_size", "max_cells", "skin", "lds_order"):
        assert g2.tuned[k] == t[k]
