"""The scene path's step loop captured as one HIP graph and replayed.

    g = StepGraph(integrator, model, dt, steps=20)   # after at least one simulate() of the model
    for k in range(20):
        command(model, k)                            # host: the model's kinematic targets
        g.stage(k)                                   # copied into the pinned slot of step k
    g.replay()                                       # the 20 steps: one graph launch

A replay holds, on the solver's stream:
* one host-to-device copy of the staged targets of all its steps, from a pinned buffer (two
  buffers and two graphs alternate, so that the host stages a replay while the last one runs);
* per step, a device copy of that step's targets into the solver's target buffer, then the
  scene kernel (d2f_step_scene), which runs the whole implicit step of every scene.

This is ImplicitIntegrator.simulate (reference implicit/integrator.py:137-149) K times over, with
the per-step host work taken out of the loop: the targets are staged in one buffer, and the
actuators' commands (pressures, tensions, activations, spring anchors, external forces) are
mirrored to the device once per replay, before its first step, as simulate() would mirror them
at the first step of the chunk. The graph captures the solver's device buffers and the step's
parameters (dt, iteration budget, options, launch shape): replay() refuses to run once the
model's solver, its layout or its parameters have changed; capture a new graph then.

Only an unrouted scene path can be captured. The per-kernel path polls the device from the host
(kinematic sweep, convergence checks), which a graph cannot hold, and a routed batch runs host
bookkeeping between its two launches.
"""
from __future__ import annotations

import numpy as np
import torch


class StepGraph:
    def __init__(self, integrator, model, dt, steps):
        if steps < 1:
            raise ValueError("steps must be >= 1")
        s = model._solver
        if s is None:
            raise RuntimeError("step the model once with simulate() before capturing its step loop")
        if s.path != "scene":
            raise ValueError("only the scene path can be captured: the per-kernel path polls the "
                             "device from the host")
        if s._route_t is not None:
            raise ValueError("this batch routes scenes to the per-kernel path; such a step cannot be "
                             "captured")
        self.integ, self.model, self.dt, self.steps = integrator, model, float(dt), int(steps)
        integrator.attach(model, dt)
        if s._layout is None:
            s.build_layout()
        self.solver = s
        self._key = self._params_key()
        n_bc = s.n_bc
        shape = (self.steps, max(n_bc, 1), 2)
        # two pinned staging buffers, one graph reading each: the host stages the next replay's
        # targets while the last one runs
        self._host = [torch.zeros(shape, dtype=torch.float64, pin_memory=True) for _ in range(2)]
        self._dev = torch.zeros(shape, dtype=torch.float64, device=s.dev)
        self._done = [None, None]          # event after the last replay that read each buffer
        self._next = 0                     # the buffer (and graph) of the next replay
        self._staged = set()
        self.graphs = []
        torch.cuda.synchronize(s.dev)
        for b in range(2):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                if n_bc:
                    self._dev.copy_(self._host[b], non_blocking=True)
                for k in range(self.steps):
                    if n_bc:
                        s.bc_target.copy_(self._dev[k, :n_bc])
                    s.step()
            self.graphs.append(g)
        torch.cuda.synchronize(s.dev)

    def _params_key(self):
        # every device pointer, size and step parameter the kernel was captured with
        s = self.solver
        return (id(s), bytes(s.scene_struct()), bytes(s.params_struct()))

    def stage(self, k, targets=None):
        """Targets of step k of the next replay: `targets` (the model's full (N, 2) target
        array) or the model's current ones. A step left unstaged takes the model's targets at
        the time of the replay."""
        if not 0 <= k < self.steps:
            raise IndexError(f"step {k} of a graph of {self.steps}")
        b = self._next
        if self._done[b] is not None:
            self._done[b].synchronize()    # the replay before last has read this buffer
            self._done[b] = None
        s = self.solver
        if s.n_bc:
            t = self.model._targets if targets is None else targets
            self._host[b][k, :s.n_bc] = torch.from_numpy(np.asarray(t, np.float64)[s.bc_idx_np])
        self._staged.add(k)

    def replay(self):
        """Run the captured steps on the current stream; model.time advances by steps * dt."""
        m, s = self.model, self.solver
        if m._solver is not s:
            raise RuntimeError("the model has a new solver since the capture; capture a new graph")
        self.integ._sync_actuation(m)
        if self._params_key() != self._key:
            raise RuntimeError("the solver's layout or step parameters changed since the capture; "
                               "capture a new graph")
        for k in set(range(self.steps)) - self._staged:
            self.stage(k)
        b = self._next
        self.graphs[b].replay()
        self._done[b] = torch.cuda.Event()
        self._done[b].record(torch.cuda.current_stream(s.dev))
        self._next, self._staged = 1 - b, set()
        m.time += self.steps * self.dt

    def run(self, command, first_step):
        """command(k) for k = first_step .. first_step + steps - 1, each staged, then one
        replay (which returns before the device has finished: the next run() stages its
        targets meanwhile)."""
        for k in range(self.steps):
            command(first_step + k)
            self.stage(k)
        self.replay()
