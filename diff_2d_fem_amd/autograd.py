"""Differentiable rollouts: T implicit steps behind a torch.autograd.Function, differentiated by
the implicit adjoint (AdjointPNCG2D) with step checkpointing and recomputation.

    roll = Rollout(model, ImplicitIntegrator(...), dt, T, every=10)
    x_T = roll(scale, pressure)          # torch tensors; x_T (N, 2) float64 on the GPU
    loss = ((x_T[tip] - target) ** 2).sum()
    loss.backward()                      # scale.grad (n_elements,), pressure.grad

scale      per-element multiplier of the elements' moduli (mu, lambda); 1 = the model's own.
           The gradient with respect to a pixel body's modulus map is model.to_pixels(handle,
           scale.grad / E_pixel) for a body of modulus E_pixel; scale.grad itself is dJ/d(ln E).
pressure   chamber pressures [Pa] in the model's chamber order: (n_chambers,) held for the whole
           rollout, or (T, n_chambers) per step (step k uses row k-1). None: the model's.

The forward pass keeps (x, v) every `every` steps. The backward pass walks the segments from
the last: it restores a segment's first state, re-runs its steps keeping every state, and
takes the damped backward-Euler adjoint recursion (reference implicit/adjoint.py:372-399)
through them, so memory holds about T / every + every states instead of T. The reference
keeps every state of the trajectory in the caller's list instead.

The model's actuators other than pressure must have a tangent in the adjoint
(AdjointPNCG2D._check_terms). Kinematic targets are set by `command(model, k)` before step k
(k = 1..T), in the forward pass and again in each recomputation.
"""
from __future__ import annotations

import numpy as np
import torch

from .adjoint import AdjointPNCG2D

F64 = torch.float64


class Rollout:
    def __init__(self, model, integrator, dt, T, every=10, command=None, adjoint_kw=None,
                 use_disc=False, use_sdf=False, use_self=False, use_ipc=None):
        if T < 1 or every < 1:
            raise ValueError("T and every must be >= 1")
        self.model, self.integ, self.dt, self.T, self.every = model, integrator, float(dt), int(T), int(every)
        self.command = command
        self.adjoint_kw = dict(rtol=1e-12, maxiter=20000) if adjoint_kw is None else dict(adjoint_kw)
        self.use = dict(use_disc=use_disc, use_sdf=use_sdf, use_self=use_self, use_ipc=use_ipc)
        self.solver = integrator.attach(model, dt)
        # a density of rho = scale with Emin = 0: moduli = scale x the model's own
        self.adj = AdjointPNCG2D(self.solver, Emin=0.0)
        self.adj._check_terms()
        self.n_ch = len(model.actuation.chambers)
        self.recomputed_steps = 0

    # ---- inputs
    def _pressures(self, pressure):
        if self.n_ch == 0 or pressure is None:
            return None
        p = pressure.detach().to("cpu", F64).numpy()
        if p.ndim == 1:
            p = np.broadcast_to(p, (self.T, self.n_ch))
        if p.shape != (self.T, self.n_ch):
            raise ValueError(f"pressure must be ({self.n_ch},) or ({self.T}, {self.n_ch}); got {p.shape}")
        return p

    def _set_scale(self, scale):
        s = scale.detach().to("cpu", F64).numpy().reshape(-1)
        if s.shape[0] != self.adj.n_tri:
            raise ValueError(f"scale has {s.shape[0]} entries, the mesh {self.adj.n_tri} elements")
        self.adj.set_density(s)

    # ---- stepping
    def _step(self, k, p):
        m = self.model
        if self.command is not None:
            self.command(m, k)
        if p is not None:
            for c in range(self.n_ch):
                m.actuation.chambers[c].pressure = float(p[k - 1, c])
            m.actuation.commanded()
        self.integ.simulate(m, self.dt)

    _AFFINE_STATE = ("ag_q", "ag_q0", "ag_pq", "ag_gq", "ag_gq_prev", "ag_zq")

    def _state(self):
        s = self.solver
        groups = tuple(getattr(s, k).clone() for k in self._AFFINE_STATE) if s.affine is not None else ()
        return s.x.clone(), s.v.clone(), self.model.time, groups

    def _restore(self, st):
        s = self.solver
        s.x.copy_(st[0])
        s.v.copy_(st[1])
        self.model.time = st[2]
        for k, t in zip(self._AFFINE_STATE, st[3]):
            getattr(s, k).copy_(t)

    def forward(self, scale, pressure):
        """T steps from the model's current state; returns x_T (N, 2)."""
        self._set_scale(scale)
        self._p = self._pressures(pressure)
        self.checkpoints = {0: self._state()}
        for k in range(1, self.T + 1):
            self._step(k, self._p)
            if k % self.every == 0 and k < self.T:
                self.checkpoints[k] = self._state()
        self.final = self._state()
        return self.solver.x.clone()

    def backward(self, grad_xT):
        """dJ/dscale (n_elements,) and dJ/dpressure (shape of the pressure input) for the
        seed dJ/dx_T."""
        adj, T = self.adj, self.T
        zero = torch.zeros_like(self.solver.x)
        lam1, lam2 = zero, zero
        g = torch.zeros(adj.n_tri, dtype=F64, device=adj.dev)
        gp = np.zeros((T, self.n_ch))
        seed = grad_xT.detach().to(adj.dev, F64).reshape(-1, 2)
        starts = sorted(self.checkpoints)
        adj.refresh_mass()
        try:
            for c0 in reversed(starts):
                c1 = min(c0 + self.every, T)
                self._restore(self.checkpoints[c0])
                xs = {}
                for k in range(c0 + 1, c1 + 1):
                    self._step(k, self._p)
                    self.recomputed_steps += 1
                    xs[k] = self.solver.x.clone()
                for k in range(c1, c0, -1):
                    if self._p is not None:
                        self.solver.set_pressure(self._p[k - 1])
                    lam, gk, gpk = adj.adjoint_step(xs[k], lam1, lam2, seed if k == T else None,
                                                    **self.use, **self.adjoint_kw)
                    g += gk
                    if gpk is not None:
                        gp[k - 1] = gpk.cpu().numpy()
                    lam1, lam2 = lam, lam1
        finally:
            self._restore(self.final)
        return g, torch.as_tensor(gp, dtype=F64)

    def __call__(self, scale, pressure=None):
        return _RolloutFn.apply(self, scale, pressure if pressure is not None else torch.zeros(0, dtype=F64))


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, roll, scale, pressure):
        ctx.roll = roll
        ctx.s_device = scale.device
        ctx.p_shape = tuple(pressure.shape)
        ctx.p_device = pressure.device
        return roll.forward(scale, pressure if pressure.numel() else None)

    @staticmethod
    def backward(ctx, grad_x):
        roll = ctx.roll
        g, gp = roll.backward(grad_x)
        if len(ctx.p_shape) == 1 and ctx.p_shape[0] > 0:
            gp = gp.sum(0)
        elif len(ctx.p_shape) == 0 or ctx.p_shape == (0,):
            gp = None
        return None, g.to(ctx.s_device), (None if gp is None else gp.to(ctx.p_device))
