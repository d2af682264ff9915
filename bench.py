"""Throughput of the batched implicit step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--path scene|plain]
                    [--precision mixed|double] [--no-plain] [--no-cpu] [--no-large]
                    [--no-adjoint] [--no-graph] [--no-tune] [--no-c5] [--no-table]
    python bench.py --table [--precision double|mixed]

The workload is the contact-heavy configuration: 256 scenes of two soft pads on kinematic
jaws gripping and lifting a rigid block (diff_2d_fem_amd.workloads.pads_lift_block). W
untimed steps bring the scenes into the gripped state, then K steps are timed, bracketed by
torch.cuda.synchronize(). A "step" is one ImplicitIntegrator.simulate of every scene,
including the host command that sets the jaws' targets.

Printed: one JSON line. `value` is scene-steps/s of the requested path and precision;
`plain_fp64` is the per-kernel float64 path on the same workload and steps (the
denominator of the batched-vs-plain ratio); `pos_err_1step` is, after the timed steps, the
node-position difference between one step of the measured precision and one float64 step
from the identical state, over each scene's block size; `roofline` is the dominant kernel
(the scene path's one step kernel, or the plain path's element kernel) with its algorithmic
bytes over its event-timed duration against HBM peak; `element_roofline` is the float64
elastic gradient+diagonal element kernel alone on this workload's mesh, and
`element_roofline_large` the same kernel at the size of the star-mesh configuration (45 M
triangles, working set far beyond the Infinity Cache); `adjoint_c4` the differentiable rollout of
the adjoint configuration (512 pneumatic fingers: forward and backward ms per step); `graph`
the scene path's K steps captured as one HIP graph and replayed (staged targets, one graph
launch) beside the eager per-step loop; `tuning` the scene kernel's plan that the autotuner
chose from the warmed-up state (--no-tune keeps the default); `c5` the large uneven batch at 256
star-mesh scenes (2k-20k triangles each; mixed precision, tuned, eager and graph-replayed, and
the per-kernel float64 path on the same steps: a batch whose scenes converge at different
iterations); `reference_table` the reference's own throughput scenes (beam, strip, gripped
block) at 1 and 4096 environments beside its published figures (--table runs every count);
`cpu_baseline` times the float64 CPU oracle on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def graph_leg(res, steps, per_replay=5):
    """The same steps as a HIP graph (diff_2d_fem_amd.graph.StepGraph): each replay holds the
    staged targets' host-to-device copy and `per_replay` steps; the host stages the next replay
    while one runs. One untimed replay, then `steps` timed steps continuing the measured
    trajectory (host clock, and HIP events around each replay on its stream)."""
    import torch
    from diff_2d_fem_amd.graph import StepGraph
    m, w, integ = res["m"], res["w"], res["integ"]
    g = StepGraph(integ, m, w["dt"], per_replay)
    step = res["step"]
    g.run(w["command"], step + 1)
    step += per_replay
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    n = max(steps // per_replay, 1)
    ev = []
    t0 = time.perf_counter()
    for r in range(n):
        for k in range(per_replay):
            w["command"](step + 1 + k)
            g.stage(k)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        g.replay()
        e1.record(st)
        ev.append((e0, e1))
        step += per_replay
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    res["step"] = step
    return dict(steps_per_replay=per_replay, steps=n * per_replay,
                ms_per_step=round(t / (n * per_replay) * 1e3, 3),
                replay_ms_per_step=round(sum(a.elapsed_time(b) for a, b in ev) / (n * per_replay), 3),
                eager_ms_per_step=round(res["t"] / res["n_steps"] * 1e3, 3),
                eager_launch_ms=round(res["launch_ms"], 3))


def c5_leg(n_scenes=256, steps=10, warmup=3, plain=True):
    """The large uneven batch (star meshes of 2k-20k triangles with affine regions, tendons and
    the ground; mixed precision) at a stated scene count: warm-up, the autotuner, `steps` eager
    steps (ms/step, PNCG iterations/s, the step kernel's roofline), then the same number of
    steps replayed as HIP graphs. Scenes of this batch converge at different iterations (the
    small meshes long before their budget): `plain_fp64` runs the same steps on the per-kernel
    float64 path, which stops when every scene has converged."""
    t0 = time.perf_counter()
    res = run_gpu("c5", n_scenes, "scene", "mixed", steps, warmup, tune=True)
    m = res["m"]
    out = dict(workload=res["w"]["name"], scenes=m.n_envs, nodes=m.n_nodes, triangles=m.n_elements,
               dt=res["w"]["dt"], setup_and_warmup_s=round(time.perf_counter() - t0 - res["t"], 1),
               ms_per_step=round(res["t"] / steps * 1e3, 3),
               scene_steps_per_s=round(m.n_envs * steps / res["t"], 1),
               pncg_iters_per_s=round(res["iters"] / res["t"], 1),
               iters_per_scene_per_step=round(res["iters"] / m.n_envs / steps, 1),
               converged_frac=round(float(np.mean(res["stats"]["converged"])), 4),
               max_iter=int(res["stats"]["max_iter"]),
               element_lds=bool(m._solver.nacc_lds),
               roofline=scene_roofline(res, "mixed"))
    if res["tuned"] is not None:
        out["tuning"] = {k: res["tuned"][k] for k in ("block_size", "max_cells", "skin", "lds_order")}
    out["graph"] = graph_leg(res, steps)
    if plain:
        pr = run_gpu("c5", n_scenes, "plain", "double", steps, warmup)
        out["plain_fp64"] = dict(ms_per_step=round(pr["t"] / steps * 1e3, 3),
                                 iters_per_scene_per_step=round(pr["iters"] / n_scenes / steps, 1),
                                 converged_frac=round(float(np.mean(pr["stats"]["converged"])), 4))
        out["speedup_vs_plain_fp64"] = round(pr["t"] / res["t"], 3)
    return out


# env-steps/s of the reference's own table (README.md:206-210; RTX A5000, float64, the card about
# 40 % loaded by another job), at 1, 16, 64, 256, 1024 and 4096 environments
TABLE_COUNTS = (1, 16, 64, 256, 1024, 4096)
TABLE_PUBLISHED = {"beam": (27, 353, 666, 909, 1037, 1061), "strip": (14, 187, 472, 678, 751, 744),
                   "grip": (9, 97, 285, 429, 531, 521)}


def env_table(precision="double", counts=TABLE_COUNTS, scenes=("beam", "strip", "grip"), tune=True):
    """The reference's throughput table on this build (protocol of tools/benchmark_envs.py:73-93):
    build the scene copied into n environments, run its lead steps (reaching the named state),
    synchronize, then time 20 steps (10 from 1024 environments on) including the host command.
    Scene path; with `tune` the autotuner picks the launch shape and LDS plan when the solver is
    built (inside the untimed lead). Per row: ms/step, iterations (mean over steps of the batch's
    largest count), host command ms/step, env-steps/s and its ratio to the published figure."""
    import torch
    from diff_2d_fem_amd import ImplicitIntegrator
    from diff_2d_fem_amd.workloads import TABLE_SCENES
    rows = []
    for name in scenes:
        for n in counts:
            t0 = time.perf_counter()
            m, kw, command, dt, lead = TABLE_SCENES[name](n)
            integ = ImplicitIntegrator(path="scene", precision=precision, autotune=tune, **kw)
            step = 0
            for _ in range(lead):
                step += 1
                command(step)
                integ.simulate(m, dt)
            torch.cuda.synchronize()
            build_s = time.perf_counter() - t0
            steps = 10 if n >= 1024 else 20
            its, t_host = [], 0.0
            t1 = time.perf_counter()
            for _ in range(steps):
                step += 1
                c0 = time.perf_counter()
                command(step)
                t_host += time.perf_counter() - c0
                integ.simulate(m, dt)
                its.append(m._solver.iters.max())              # on the device: read after the loop
            torch.cuda.synchronize()
            t = (time.perf_counter() - t1) / steps
            it = float(np.mean([int(v) for v in its]))
            st = integ.stats(m)
            pub = TABLE_PUBLISHED[name][TABLE_COUNTS.index(n)] if n in TABLE_COUNTS else None
            row = dict(scene=name, envs=n, nodes=m.n_nodes, ms_per_step=round(t * 1e3, 3), iterations=round(it, 1),
                       max_iter=int(st["max_iter"]), host_ms_per_step=round(t_host / steps * 1e3, 3),
                       env_steps_per_s=round(n / t, 1), ns_per_node_iteration=round(t / it / m.n_nodes * 1e9, 3),
                       converged=int(np.sum(st["converged"])), lead_and_setup_s=round(build_s, 1),
                       block_size=int(m._solver.block_size),
                       published_a5000=pub, ratio=None if pub is None else round(n / t / pub, 1))
            rows.append(row)
            print(f"table {name} {n} envs: {row['ms_per_step']} ms/step, {row['env_steps_per_s']} env-steps/s",
                  file=sys.stderr, flush=True)
            del m, integ
    return rows


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--envs", type=int, default=None)
    ap.add_argument("--path", default=None)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-large", action="store_true", help="skip the 45 M-triangle element-kernel leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 512-finger adjoint leg")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay leg")
    ap.add_argument("--no-tune", action="store_true", help="scene path: keep the default plan")
    ap.add_argument("--no-c5", action="store_true", help="skip the 256-scene star-mesh leg")
    ap.add_argument("--table", action="store_true",
                    help="only the reference's throughput table (beam, strip, grip at 1-4096 environments)")
    ap.add_argument("--no-table", action="store_true",
                    help="skip the reference table's rows at 1 and 4096 environments (mixed, tuned)")
    return ap.parse_args()


def build(config, n_env):
    from diff_2d_fem_amd import workloads as W
    fn = W.CONFIGS[config]
    kw = {}
    if config == "c5":
        kw["progress"] = lambda k, n, what: print(f"c5 set-up: {k}/{n} scenes, {what}", file=sys.stderr, flush=True)
    return fn(**kw) if n_env is None else fn(n_env, **kw)


def run_gpu(config, envs, path, precision, steps, warmup, tune=False):
    """W warm-up steps, the autotuner (scene path, when asked: it measures its candidates from
    the warmed-up state and puts the state back), then K timed steps. HIP events recorded on the
    launch stream right around each solver launch give the step kernel's own duration."""
    import torch
    from diff_2d_fem_amd import ImplicitIntegrator
    w = build(config, envs)
    m = w["model"]
    integ = ImplicitIntegrator(path=path, precision=precision)
    step = 0
    for _ in range(warmup):
        step += 1
        w["command"](step)
        integ.simulate(m, w["dt"])
        if config == "c5":                          # minutes per step at thousands of scenes
            torch.cuda.synchronize()
            print(f"c5 {path} {precision}: warm-up step {step} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    tuned = None
    if tune and path == "scene":
        from diff_2d_fem_amd.tune import autotune
        tuned = autotune(integ, m, w["dt"], cache=False)
    s = m._solver
    events = []
    launch = s.step

    def timed_launch():
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        launch()
        e1.record(st)
        events.append((e0, e1))
    acc = torch.zeros(m.n_envs, dtype=torch.int64, device="cuda")
    s.step = timed_launch
    t0 = time.perf_counter()
    for _ in range(steps):
        step += 1
        w["command"](step)
        integ.simulate(m, w["dt"])
        acc += s.iters                              # on the stream: no host sync in the loop
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    del s.step
    launch_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in events]))
    st = integ.stats(m)
    iters = int(acc.sum().item())
    return dict(t=t, m=m, w=w, integ=integ, stats=st, iters=iters, step=step, launch_ms=launch_ms,
                n_steps=steps, tuned=tuned)


def _snapshot(s):
    import torch
    return {k: v.clone() for k, v in vars(s).items() if torch.is_tensor(v)}


def _restore(s, snap):
    for k, v in snap.items():
        getattr(s, k).copy_(v)


def one_step_error(res):
    """Position error of one step of the measured precision against one float64 step of the
    same path from the identical state (same commands, float64 defaults: eps 1e-10), over
    each scene's block size (or the workload's body size)."""
    import torch
    m, w, s = res["m"], res["w"], res["m"]._solver
    snap, t_model = _snapshot(s), m.time
    prec, eps = s.precision, s.eps
    k = res["step"] + 1
    w["command"](k)
    res["integ"].simulate(m, w["dt"])
    x_meas, it_meas = s.get_x(), s.iters.cpu().numpy().copy()
    _restore(s, snap)
    m.time = t_model
    s.precision, s.eps = "double", 1e-10
    w["command"](k)
    res["integ"].simulate(m, w["dt"])
    torch.cuda.synchronize()
    x_ref, it_ref = s.get_x(), s.iters.cpu().numpy().copy()
    _restore(s, snap)
    m.time = t_model
    s.precision, s.eps = prec, eps
    env = np.asarray(m.me.node2env)
    err = np.linalg.norm(x_meas - x_ref, axis=1)
    per_env = np.zeros(m.n_envs)
    np.maximum.at(per_env, env, err)
    size = np.asarray(w.get("sides", np.full(m.n_envs, w["size"])), np.float64)
    rel = per_env / size
    return dict(max=float(rel.max()), mean=float(rel.mean()), iters_measured=float(it_meas.mean()),
                iters_fp64=float(it_ref.mean()), reference="float64 " + s.path + " path, eps 1e-10",
                step=k)


def measured_traffic(kernel, res):
    """HBM bytes per launch of `kernel` on this workload from the committed rocprofv3 PMC
    summary (profiles/c3_traffic.json: FETCH_SIZE and WRITE_SIZE scaled by a copy kernel's
    readings), or None when there is no measurement for it."""
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "c3_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("workload") != res["w"]["name"] or d.get("scenes") != res["m"].n_envs:
        return None
    k = d["kernels"].get(kernel)
    return None if k is None else int(k["traffic_bytes"])


def scene_roofline(res, precision):
    """The step kernel of the scene path (one launch per step): algorithmic bytes per launch
    = sum over scenes of iterations x per-iteration bytes, where one iteration reads each
    scene's element constants twice (gradient+diagonal and curvature: T (12 + 6 s_e) for the
    three node indices, D_m^-1 and the area-weighted moduli) and touches node state for 240
    bytes per node (the per-kernel passes' 272 less the 32 of copying the gradient into the
    previous-gradient array, which the kernel replaces by swapping the two arrays' roles each
    iteration); s_e = 4 in mixed, 8 in float64."""
    s = res["m"]._solver
    se = 4 if precision == "mixed" else 8
    it_per_env = res["iters"] / res["m"].n_envs / res["n_steps"]
    bytes_ = it_per_env * (2.0 * (12 + 6 * se) * s.n_tri + 240.0 * s.N)
    dur = res["launch_ms"] * 1e-3
    ach = bytes_ / dur / 1e9
    # the kernel's lean variant (no tendons, springs, pressure, rigid modes, node contact or line
    # search) or its full one: the traffic file names the one the workload ran
    base = f"d2f::k_scene<{'float' if precision == 'mixed' else 'double'}, {s.block_size}, "
    traffic = measured_traffic(base + "false>", res) or measured_traffic(base + "true>", res)
    return dict(kernel=f"k_scene ({precision})", bound="hbm", achieved=round(ach, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                traffic=traffic,
                avg_us=round(dur * 1e6, 1), bytes_per_launch=int(bytes_),
                iters_per_scene_per_step=round(it_per_env, 2))


def elastic_roofline(res):
    """Time the float64 elastic gradient+diagonal kernel (tiled, with the solver's node
    ownership) on the run's final state with HIP events on the launch stream; algorithmic
    bytes = 68 T + 48 N per launch."""
    import ctypes as C
    import torch
    from diff_2d_fem_amd import _lib
    s = res["m"]._solver
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    h = C.c_void_p(stream.cuda_stream)
    g = torch.zeros_like(s.x)
    d = torch.zeros_like(s.x)
    P = lambda t: C.c_void_p(t.data_ptr())
    tiles = s.elem_tiles_struct()
    args = (P(s.x), P(s.tri), P(s.Bm), P(s.area), P(s.mu), P(s.lam), s.n_tri, s.h2, 1, C.byref(tiles),
            P(g), P(d), h)
    for _ in range(3):
        L.d2f_elastic_grad_diag_f64(*args)
    n = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n):
        L.d2f_elastic_grad_diag_f64(*args)
    e1.record(stream)
    torch.cuda.synchronize()
    dur = e0.elapsed_time(e1) / n * 1e-3
    bytes_ = 68.0 * s.n_tri + 48.0 * s.N
    ach = bytes_ / dur / 1e9
    return dict(kernel="k_elastic_grad_diag (float64)", bound="hbm", achieved=round(ach, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                traffic=measured_traffic("d2f::k_elastic_grad_diag_tiled<false>", res),
                avg_us=round(dur * 1e6, 2), bytes_per_launch=int(bytes_))


def element_roofline_large(n_meshes=32, copies=128, seed=0):
    """The float64 element kernel at the size of the star-mesh configuration (about 45 M
    triangles, 23 M nodes, a ~3 GB working set far beyond the 256 MB Infinity Cache): 32
    seeded Delaunay star meshes of 2k-20k triangles (diff_2d_fem_amd.workloads) tiled 128
    times, randomly deformed, with the element kernel's node tiles. HIP events on the
    launch stream; algorithmic bytes = 68 T + 48 N per launch."""
    import ctypes as C
    import torch
    from diff_2d_fem_amd import _lib
    from diff_2d_fem_amd.layout import element_tiles, tiles_struct
    from diff_2d_fem_amd.mesh import lame
    from diff_2d_fem_amd.workloads import _star_mesh
    Vs, Ts, Bs, As, MUs, LAs = [], [], [], [], [], []
    for i in range(n_meshes):
        rng = np.random.default_rng([seed, i])
        V, T, _ = _star_mesh(rng, 10e-3, int(rng.integers(2000, 20001)))
        D = np.stack([V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]]], 2)     # columns = edges
        mu, la = lame(float(np.exp(rng.uniform(np.log(2e5), np.log(2e6)))), float(rng.uniform(0.40, 0.45)))
        Vs.append(V + rng.normal(scale=2e-5, size=V.shape))
        Ts.append(T)
        Bs.append(np.linalg.inv(D).reshape(-1, 4))
        As.append(0.5 * np.abs(np.linalg.det(D)))
        MUs.append(np.full(len(T), mu))
        LAs.append(np.full(len(T), la))
    nv = np.array([len(v) for v in Vs])
    off = np.concatenate([[0], np.cumsum(np.tile(nv, copies))[:-1]])
    tri = np.concatenate([Ts[k % n_meshes] + off[k] for k in range(n_meshes * copies)]).astype(np.int32)
    x = np.concatenate([Vs[k % n_meshes] for k in range(n_meshes * copies)])
    cat = lambda L: np.concatenate([L[k % n_meshes] for k in range(n_meshes * copies)])
    Bm, area, mu, lam = cat(Bs), cat(As), cat(MUs), cat(LAs)
    T, N = len(tri), len(x)
    dev = torch.device("cuda")
    et = element_tiles(tri, N, device=dev)
    tiles = tiles_struct(et)
    tt = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=dev)
    xd, trid, Bd, Ad, mud, lad = tt(x), tt(tri, torch.int32), tt(Bm), tt(area), tt(mu), tt(lam)
    g = torch.zeros_like(xd)
    d = torch.zeros_like(xd)
    L = _lib.lib()
    st = torch.cuda.current_stream()
    h = C.c_void_p(st.cuda_stream)
    P = lambda t: C.c_void_p(t.data_ptr())
    args = (P(xd), P(trid), P(Bd), P(Ad), P(mud), P(lad), T, 1e-6, 1, C.byref(tiles), P(g), P(d), h)
    for _ in range(3):
        _lib.check(L.d2f_elastic_grad_diag_f64(*args), "elastic grad/diag")
    n = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(n):
        L.d2f_elastic_grad_diag_f64(*args)
    e1.record(st)
    torch.cuda.synchronize()
    dur = e0.elapsed_time(e1) / n * 1e-3
    bytes_ = 68.0 * T + 48.0 * N
    ach = bytes_ / dur / 1e9
    return dict(kernel="k_elastic_grad_diag (float64)", workload=f"{n_meshes} star meshes x {copies}",
                triangles=int(T), nodes=int(N), bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4), avg_us=round(dur * 1e6, 1),
                bytes_per_launch=int(bytes_), max_chunk_nodes=et["max_nodes"],
                owned_frac=round(float(et["owned"].float().mean()), 4),
                traffic=large_traffic(f"{n_meshes} star meshes x {copies}", int(T)))


def large_traffic(workload, T):
    """HBM bytes per launch of the element entry point at the large size (its chunk pass and
    its halo pass) from the committed rocprofv3 PMC summary profiles/element_large_traffic.json,
    or None when there is no measurement for this workload."""
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "element_large_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload or d.get("triangles") != T:
        return None
    return int(sum(k["traffic_bytes"] for k in d["kernels"].values()))


def adjoint_leg(n_fingers=512, T=10, every=5):
    """The adjoint configuration: 512 pneumatic fingers (diff_2d_fem_amd.workloads,
    pneumatic_fingers), loss |tip - target|^2 after T steps of dt 2 ms, gradients with respect
    to every chamber pressure and every element's modulus scale through the autograd rollout
    (checkpoints every `every` steps, recomputation in the backward pass, float64, per-scene
    device MINRES at the reference's rtol 1e-9). Times the forward pass and the backward pass
    (recomputation + one adjoint solve per step) with synchronisation on both sides, after an
    untimed 2-step rollout from the same initial state."""
    import torch
    from diff_2d_fem_amd import ImplicitIntegrator
    from diff_2d_fem_amd import workloads as W
    from diff_2d_fem_amd.autograd import Rollout
    w = W.pneumatic_fingers(n_fingers)
    m = w["model"]
    tips = torch.as_tensor([int(m.body_nodes(h)[np.argmax(m.rest_x(h)[:, 0])]) for h in w["bodies"]],
                           device="cuda")
    targets = torch.as_tensor(w["targets"], device="cuda")
    integ = ImplicitIntegrator(path="scene", precision="double", self_collision=False)
    # untimed warm-up: a 2-step rollout through its forward and backward passes builds the scene
    # layout, loads the kernels and sizes the allocator; then the model goes back to rest
    warm = Rollout(m, integ, w["dt"], 2, every=1, adjoint_kw=dict(rtol=1e-9, maxiter=5000))
    ws_ = torch.ones(m.n_elements, dtype=torch.float64, requires_grad=True)
    wp_ = torch.tensor([c.pressure for c in m.actuation.chambers], requires_grad=True)
    ((warm(ws_, wp_)[tips] - targets) ** 2).sum().backward()
    del warm
    m.reset()
    roll = Rollout(m, integ, w["dt"], T, every=every, adjoint_kw=dict(rtol=1e-9, maxiter=5000))
    scale = torch.ones(m.n_elements, dtype=torch.float64, requires_grad=True)
    p = torch.tensor([c.pressure for c in m.actuation.chambers], requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = ((roll(scale, p)[tips] - targets) ** 2).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    it = integ.stats(m)["iters"]
    return dict(workload="pneumatic_fingers", scenes=n_fingers, nodes=m.n_nodes, triangles=m.n_elements,
                chambers=len(m.actuation.chambers), steps=T, dt=w["dt"], checkpoint_every=every,
                forward_ms_per_step=round((t1 - t0) / T * 1e3, 3),
                backward_ms_per_step=round((t2 - t1) / T * 1e3, 3),
                recomputed_steps=roll.recomputed_steps, pncg_iters_last_step=int(it.max()),
                minres_iters_last_solve=int(getattr(roll.adj, "minres_iters", 0)),
                minres_unconverged=int(getattr(roll.adj, "minres_fail", 0)),
                grad_norm_scale=float(scale.grad.norm()), grad_norm_pressure=float(p.grad.norm()))


def _compiled_rate(config, n_scenes, warmup, steps, threads):
    """Scene-steps/s of the compiled oracle step (oracle/cpu_step.cpp) on the first n_scenes
    scenes of the workload: `warmup` untimed steps, then `steps` timed."""
    from oracle.cpu import CpuStepper
    from oracle.pncg import OracleIntegrator
    from diff_2d_fem_amd import workloads as W
    w = W.CONFIGS[config](n_scenes)
    m = w["model"]
    cs = CpuStepper(OracleIntegrator(), m, w["dt"], threads=threads)
    for k in range(1, warmup + 1):
        w["command"](k)
        cs.simulate(m, w["dt"])
    t = 0.0
    for k in range(warmup + 1, warmup + steps + 1):
        w["command"](k)
        t0 = time.perf_counter()
        cs.simulate(m, w["dt"])
        t += time.perf_counter() - t0
    return n_scenes * steps / t, t


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, steps):
    """CPU baselines on the host cores of this run. The compiled restatement of the reference
    algorithm (oracle/cpu_step.cpp: float64, one environment after another, kernel loops in
    index order) timed over the same step range as the GPU's timed steps (after args.warmup
    untimed steps): with OpenMP over environments on every core this process may use (the
    reported value), and serially on one core. Also the float64 numpy oracle on one scene from
    its first step, whose positions serve oracle_error. Returns (entry, numpy oracle positions)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import cpu
    from oracle.pncg import OracleIntegrator
    from diff_2d_fem_amd import workloads as W
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    cpu.build()
    tsteps = 3
    par, tp = _compiled_rate(args.config, 4 * cores, args.warmup, tsteps, cores)
    ser, ts = _compiled_rate(args.config, 8, args.warmup, tsteps, 1)
    w = W.CONFIGS[args.config](1)
    m = w["model"]
    oi = OracleIntegrator()
    xs = []
    t = 0.0
    for k in range(1, steps + 1):
        w["command"](k)
        t0 = time.perf_counter()
        oi.simulate(m, w["dt"])
        t += time.perf_counter() - t0
        xs.append(oi.solver.x.copy())
    rng = f"steps {args.warmup + 1}-{args.warmup + tsteps}"
    return dict(value=round(par, 3), unit="scene-steps/s", cores=cores, kind="port",
                sample=f"compiled float64 restatement (oracle/cpu_step.cpp), OpenMP over {4 * cores} scenes of "
                       f"{args.config}, {rng} ({tp:.1f} s), {_cpu_model()}",
                serial=dict(value=round(ser, 4), cores=1,
                            sample=f"the same, serial, 8 scenes, {rng} ({ts:.1f} s)"),
                numpy_oracle=dict(value=round(steps / t, 4), cores=1,
                                  sample=f"float64 numpy oracle, 1 scene, steps 1-{steps} ({t:.1f} s)")), xs


def oracle_error(args, xs, path, precision):
    """Position error of the GPU path against the float64 oracle on the oracle's scene (scene 0
    of the workload, the same seeded scene), step by step from rest over the oracle's steps:
    max |x_gpu - x_oracle| over the nodes, over the scene's body size."""
    from diff_2d_fem_amd import ImplicitIntegrator
    from diff_2d_fem_amd import workloads as W
    w = W.CONFIGS[args.config](1)
    m = w["model"]
    integ = ImplicitIntegrator(path=path, precision=precision)
    size = float(np.asarray(w.get("sides", [w.get("size", 1.0)]), np.float64).reshape(-1)[0])
    errs = []
    for k, xo in enumerate(xs, start=1):
        w["command"](k)
        integ.simulate(m, w["dt"])
        errs.append(float(np.linalg.norm(m.get_x() - xo, axis=1).max()) / size)
    return dict(max=max(errs), per_step=[float(f"{e:.3e}") for e in errs], steps=len(xs),
                reference="float64 numpy oracle, same scene, same commands, from rest",
                path=path, precision=precision, bound=1e-5)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); "
              f"environments are independent, so a run on N GPUs would give each GPU its own scenes "
              f"with no communication", file=sys.stderr)
        sys.exit(2)
    import torch
    from diff_2d_fem_amd import _lib
    if args.table:
        prec = args.precision or "double"
        print(json.dumps({"table": env_table(prec, tune=not args.no_tune), "precision": prec, "path": "scene",
                          "tuned": not args.no_tune,
                          "published": "README.md:206-210, RTX A5000, float64"}), flush=True)
        return
    has_scene = hasattr(_lib.lib(), "d2f_step_scene")
    path = args.path or ("scene" if has_scene else "plain")
    precision = args.precision or ("mixed" if path == "scene" else "double")
    main_run = run_gpu(args.config, args.envs, path, precision, args.steps, args.warmup,
                       tune=not args.no_tune)
    m = main_run["m"]
    B = m.n_envs
    value = B * args.steps / main_run["t"]
    out = {
        "metric": "scene-steps/s batched mixed vs plain FP64 at fixed error; % HBM peak",
        "value": round(value, 2), "unit": "scene-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(main_run["t"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": main_run["w"]["name"], "scenes": B, "nodes": m.n_nodes,
                   "triangles": m.n_elements, "dt": main_run["w"]["dt"], "path": path,
                   "precision": precision, "parallelism": "one workgroup per scene" if path == "scene"
                   else "one launch per kernel family"},
        "pncg_iters_per_s": round(main_run["iters"] / main_run["t"], 1),
        "max_iter": int(main_run["stats"]["max_iter"]),
        "converged_frac": round(float(np.mean(main_run["stats"]["converged"])), 4),
    }
    if not args.no_adjoint:
        import gc
        gc.collect()
        out["adjoint_c4"] = adjoint_leg()
    tuned = main_run["tuned"]
    if tuned is not None:
        out["tuning"] = {k: tuned[k] for k in ("block_size", "max_cells", "skin", "lds_order", "nacc_lds",
                                               "xc_lds", "excl_bits_lds", "max_pairs", "cached")}
        out["tuning"]["trials_ms_per_step"] = [t["ms"] for t in tuned["trials"]]
    if path == "scene":
        out["roofline"] = scene_roofline(main_run, precision)
    out["element_roofline"] = elastic_roofline(main_run)
    if not args.no_large:
        out["element_roofline_large"] = element_roofline_large()
    if path == "plain":
        out["roofline"] = dict(out["element_roofline"])
    if not (path == "scene" and precision == "double"):
        out["pos_err_1step"] = one_step_error(main_run)
    if path == "scene" and not args.no_graph:
        out["graph"] = graph_leg(main_run, args.steps)
    if not args.no_plain and not (path == "plain" and precision == "double"):
        plain = run_gpu(args.config, args.envs, "plain", "double", args.steps, args.warmup)
        pv = B * args.steps / plain["t"]
        out["plain_fp64"] = dict(value=round(pv, 2), ms_per_step=round(plain["t"] / args.steps * 1e3, 3))
        out["speedup_vs_plain_fp64"] = round(value / pv, 3)
    if not args.no_c5:
        out["c5"] = c5_leg(plain=not args.no_plain)
    if not args.no_table:
        rows = env_table("mixed", counts=(1, 4096))
        out["reference_table"] = dict(
            precision="mixed", published="README.md:206-210, RTX A5000, float64",
            rows=[{k: r[k] for k in ("scene", "envs", "ms_per_step", "iterations", "env_steps_per_s",
                                      "published_a5000", "ratio")} for r in rows])
    if not args.no_cpu:
        out["cpu_baseline"], xs = cpu_baseline(args, args.cpu_steps)
        out["pos_err_vs_oracle"] = oracle_error(args, xs, path, precision)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
