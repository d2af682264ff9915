"""The oracle's adjoint of the implicit step — TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the reference's AdjointPNCG2D (implicit/adjoint.py:181-399)
over an OracleSolver: the operator H = M + h^2 (K_elastic + K_colliders [+ K_node with
use_self]) restricted to the free degrees of freedom, its Jacobi preconditioner, scipy's MINRES, dJ/drho per element and
the damped backward-Euler recursion over a trajectory. The product computes the same on the
GPU (diff_2d_fem_amd.adjoint); this is its checker.

Beyond the reference (whose adjoint refuses pressure chambers): with a Pressure term on the
solver, its Hessian joins the operator (oracle.bodies.Pressure.hvp) and the recursion also
accumulates dJ/dp per chamber and step (Pressure.dp) into self.dJdp_steps. Pinned by the
FD checks of tests/test_oracle.py (Pressure.hvp against differences of its gradient,
Pressure.dp against differences in p) and by central differences of whole oracle rollouts.
"""
from __future__ import annotations

import numpy as np

from . import terms as K


class OracleAdjoint:
    def __init__(self, solver, me, Emin=1e-3):
        self.s = solver
        self.me = me
        self.Emin = float(Emin)
        self.mu0 = np.asarray(me.mu, np.float64).copy()
        self.lam0 = np.asarray(me.lam, np.float64).copy()
        self.tri = np.asarray(me.tri, np.int64)
        self.B = np.asarray(me.D_m_inv, np.float64).reshape(-1, 2, 2)
        self.area = np.asarray(me.rest_area, np.float64)
        self.N, self.n_tri = solver.N, len(self.area)
        self.h2 = float(solver.h2)
        self.free = np.asarray(solver.free, np.float64).reshape(self.N, -1)
        if self.free.shape[1] == 1:
            self.free = np.repeat(self.free, 2, axis=1)
        self.mass = np.asarray(solver.mass, np.float64)
        from .bodies import Pressure
        self.pressure = next((t for t in getattr(solver, "terms", []) if isinstance(t, Pressure)), None)

    # reference implicit/adjoint.py:41-50
    def set_density(self, rho):
        E = self.Emin + np.asarray(rho, np.float64) * (1.0 - self.Emin)
        self.s.mu = E * self.mu0
        self.s.lam = E * self.lam0

    def _cols(self, use_disc, use_sdf):
        return [c for c in self.s.colliders if (use_disc and c["kind"] == 0) or (use_sdf and c["kind"] != 0)]

    # reference implicit/adjoint.py:231-256
    def apply_H(self, x, v, use_disc=False, use_sdf=False, include_mass=True, use_self=False):
        y = np.zeros((self.N, 2))
        K.elastic_hvp(x, v, self.tri, self.B, self.area, self.s.mu, self.s.lam, self.h2, y)
        for c in self._cols(use_disc, use_sdf):
            K.collider_hvp(c, x, v, self.s.bnd, self.s.dhat, self.s.kappa, self.h2, y)
        if use_self and self.s.node is not None:
            self.s.node.hvp(x, v, self.h2, y)
        if self.pressure is not None:
            self.pressure.hvp(x, v, self.h2, y)
        return self.mass[:, None] * v + y if include_mass else y

    # reference implicit/adjoint.py:356-372
    def diag_precond(self, x, use_disc=False, use_sdf=False, use_self=False):
        g = np.zeros((self.N, 2))
        d = np.repeat(self.mass[:, None], 2, axis=1).copy()
        K.elastic_grad_diag(x, self.tri, self.B, self.area, self.s.mu, self.s.lam, self.h2, 1, g, d)
        for c in self._cols(use_disc, use_sdf):
            K.collider_grad_diag(c, x, self.s.bnd, self.s.dhat, self.s.kappa, self.h2, 1, g, d)
        if use_self and self.s.node is not None:
            self.s.node.grad_diag(x, self.h2, 1, g, d)
        return np.maximum(d.reshape(-1), 1e-30)

    # reference implicit/adjoint.py:375-404
    def adjoint_solve(self, x, rhs, use_disc=False, use_sdf=False, rtol=1e-12, maxiter=20000,
                      use_self=False):
        from scipy.sparse.linalg import LinearOperator, minres
        ff = self.free.reshape(-1) > 0.5
        nf = int(ff.sum())
        dinv = 1.0 / self.diag_precond(x, use_disc, use_sdf, use_self)[ff]

        def mv(vf):
            full = np.zeros((self.N, 2))
            full.reshape(-1)[ff] = vf
            return (self.apply_H(x, full, use_disc, use_sdf, use_self=use_self) * self.free).reshape(-1)[ff]
        A = LinearOperator((nf, nf), matvec=mv)
        M = LinearOperator((nf, nf), matvec=lambda v: dinv * v)
        lf, info = minres(A, np.asarray(rhs).reshape(-1)[ff], M=M, rtol=rtol, maxiter=maxiter)
        self.info = info
        lam = np.zeros((self.N, 2))
        lam.reshape(-1)[ff] = lf
        return lam

    # reference implicit/adjoint.py:165-178
    def drho(self, x, lam):
        F = K.ds_from_nodes(x, self.tri) @ self.B
        HH = self.area[:, None, None] * (K.nh_P(F, self.mu0, self.lam0) @ np.swapaxes(self.B, -1, -2))
        g1, g2 = HH[:, :, 0], HH[:, :, 1]
        l0, l1, l2 = lam[self.tri[:, 0]], lam[self.tri[:, 1]], lam[self.tri[:, 2]]
        s = (l0 * -(g1 + g2)).sum(1) + (l1 * g1).sum(1) + (l2 * g2).sum(1)
        return -(1.0 - self.Emin) * self.h2 * s

    # reference implicit/adjoint.py:373-399
    def transient_adjoint(self, xs, seed, drag, use_disc=False, use_sdf=False, use_self=False, **kw):
        T = len(xs) - 1
        m2 = self.mass[:, None]
        lam = {T + 1: np.zeros((self.N, 2)), T + 2: np.zeros((self.N, 2))}
        g = np.zeros(self.n_tri)
        self.dJdp_steps = None if self.pressure is None else np.zeros((T, len(self.pressure.p)))
        for k in range(T, 0, -1):
            rhs = (1.0 + drag) * m2 * lam[k + 1] - drag * m2 * lam[k + 2]
            if k in seed:
                rhs = rhs + seed[k]
            lam[k] = self.adjoint_solve(xs[k], rhs, use_disc, use_sdf, use_self=use_self, **kw)
            g += self.drho(xs[k], lam[k])
            if self.pressure is not None:
                self.dJdp_steps[k - 1] = self.pressure.dp(xs[k], lam[k], self.h2)
        return g
