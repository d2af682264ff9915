"""Autotuning of the scene kernel per workload: threads per workgroup, grid cells per scene,
pair-list margin and the order in which the optional LDS parts are granted.

    from diff_2d_fem_amd.tune import autotune
    plan = autotune(integrator, model, dt)       # or ImplicitIntegrator(..., autotune=True)

None of the knobs changes what a step computes: the grid and the pair list are exact
broadphases whatever their sizes, and the workgroup size changes only the order of the
reductions. They change how fast a scene's workgroup runs: more waves per scene against more
registers per wave, fewer cells to scan against more candidates per cell, a wider margin against
more pairs per pass, and which of the element accumulators, the positions, the exclusion bitset
and the pair list live in LDS when not all of them fit.

The search is coordinate descent over those four axes, from the solver's current plan: each
candidate steps the model from one saved state, one untimed step and `steps` timed ones (HIP
events around the step launch on its stream), and the state is put back afterwards. A candidate
that would send a scene to the per-kernel path is skipped. The winner is applied and cached in
a JSON file keyed by the device, the library's ABI version and the workload's shape (sizes,
terms, precision, iteration budget), so that the next run of the same workload applies it
without measuring. The cache file is D2FEM_TUNE_CACHE, or ~/.cache/diff_2d_fem_amd/tune.json.
"""
from __future__ import annotations

import hashlib
import json
import os
import warnings
from pathlib import Path

import torch

from . import _lib

SPACE = dict(block_size=(256, 512, 1024),
             max_cells=(256, 512, 1024, 2048),
             skin=(0.25, 0.5, 1.0),
             lds_order=(("nacc", "xc", "bits", "pairs"),
                        ("nacc", "bits", "pairs", "xc"),
                        ("bits", "pairs", "nacc", "xc")))
KNOBS = ("block_size", "max_cells", "skin", "lds_order")


def cache_path():
    p = os.environ.get("D2FEM_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "diff_2d_fem_amd" / "tune.json"


def workload_key(solver):
    """Device, library version and the workload's shape: what makes a plan transferable."""
    s = solver
    dev = torch.cuda.get_device_properties(s.dev)
    shape = dict(device=dev.name, arch=getattr(dev, "gcnArchName", ""), abi=_lib.lib().d2f_abi_version(),
                 B=s.B, N=s.N, T=s.n_tri, n_b=s.n_b, n_bc=s.n_bc, max_nb=s.max_nb, max_nn=s.max_nn,
                 precision=s.precision, iter_max=s.iter_max, conv_check=s.conv_check,
                 line_search=bool(s.line_search), ipc=s.ipc is not None, node=s.node is not None,
                 affine=0 if s.affine is None else s.affine["n"],
                 rigid=0 if s.rigid is None else s.rigid["n"],
                 pressure=0 if s.pressure is None else s.pressure["n_pe"],
                 tendon=0 if s.tendon is None else s.tendon["n_td"],
                 springs=0 if s.springs is None else s.springs["n"],
                 colliders=len(s.colliders))
    text = json.dumps(shape, sort_keys=True)
    return hashlib.sha1(text.encode()).hexdigest()[:16], shape


def _load(path):
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _save(path, table):
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError as e:
        warnings.warn(f"could not write the tuning cache {path}: {e}; the plan is applied but not kept",
                      stacklevel=3)


def _knobs(solver):
    return dict(block_size=solver.block_size, max_cells=solver.max_cells, skin=solver.skin,
                lds_order=tuple(solver.lds_order))


def _apply(solver, knobs):
    solver.set_plan(block_size=knobs["block_size"], max_cells=knobs["max_cells"], skin=knobs["skin"],
                    lds_order=tuple(knobs["lds_order"]))


def _snapshot(solver):
    return {k: v.clone() for k, v in vars(solver).items() if torch.is_tensor(v)}


def _restore(solver, snap):
    for k, v in snap.items():
        getattr(solver, k).copy_(v)


def _time_steps(solver, steps):
    """ms per step of the launch alone: one untimed step, then `steps` between HIP events."""
    st = torch.cuda.current_stream(solver.dev)
    solver.step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        solver.step()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


def autotune(integrator, model, dt, steps=2, space=None, cache=True, force=False):
    """Choose and apply the scene kernel's plan for this model; returns the plan with the
    measured ms per step of every candidate tried (empty when it came from the cache)."""
    s = integrator.attach(model, dt)
    if s.path != "scene":
        raise ValueError("the autotuner tunes the scene path; this solver takes the per-kernel path")
    if s._layout is None:
        s.build_layout()
    key, shape = workload_key(s)
    path = cache_path()
    table = _load(path) if cache else {}
    if cache and not force and key in table:
        _apply(s, table[key]["knobs"])
        return dict(s.plan(), key=key, cached=True, trials=[])
    space = dict(SPACE, **(space or {}))
    base_routed = len(s.plain_envs)
    snap, t_model = _snapshot(s), model.time
    s.set_bc_targets(model._targets)
    best = _knobs(s)
    trials, seen = [], {}

    def measure(knobs):
        _apply(s, knobs)
        if len(s.plain_envs) > base_routed:          # this plan would demote a scene
            return None
        # the effective plan: orders that grant the same LDS parts are one candidate
        plan = tuple(sorted((k, v) for k, v in s.plan().items() if k != "lds_order"))
        if plan not in seen:
            _restore(s, snap)
            seen[plan] = _time_steps(s, steps)
        trials.append(dict({k: (list(v) if k == "lds_order" else v) for k, v in knobs.items()},
                           ms=round(seen[plan], 4)))
        return seen[plan]

    try:
        best_ms = measure(best)
        for axis in KNOBS:
            for v in space[axis]:
                cand = dict(best, **{axis: tuple(v) if axis == "lds_order" else v})
                if cand == best:
                    continue
                ms = measure(cand)
                if ms is not None and ms < best_ms:
                    best, best_ms = cand, ms
    finally:
        _apply(s, best)
        _restore(s, snap)
        model.time = t_model
    if cache:
        table = _load(path)
        table[key] = dict(knobs={k: (list(v) if k == "lds_order" else v) for k, v in best.items()},
                          ms_per_step=round(best_ms, 4), shape=shape)
        _save(path, table)
    return dict(s.plan(), key=key, cached=False, ms_per_step=best_ms, trials=trials)
