"""The oracle's adjoint of the implicit step — TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the reference's AdjointPNCG2D (implicit/adjoint.py:181-399)
over an OracleSolver: the operator H = M + h^2 (K_elastic + K_colliders [+ K_node with
use_self]) restricted to the free degrees of freedom, its Jacobi preconditioner, scipy's MINRES,
dJ/drho per element and the damped backward-Euler recursion over a trajectory. The product
computes the same on the GPU (diff_2d_fem_amd.adjoint); this is its checker.

Beyond the reference, whose adjoint refuses these terms:
* Pressure chambers: their Hessian joins the operator (oracle.bodies.Pressure.hvp) and the
  recursion also accumulates dJ/dp per chamber and step (Pressure.dp) into self.dJdp_steps.
* IPC contact between bodies: the barrier's Hessian with the terms frozen at x
  (oracle.ipc.IPCContact.hvp); friction has no tangent, as for the colliders.
* Affine bodies: the slaved nodes follow x_i = J_i q, so the step's optimality conditions are
  those of the energy in (free soft nodes, q); the solve is P^T H P lam_r = P^T rhs with P that
  map, and lam = P lam_r. The regions' own elements carry their moduli in H, so at the slaved
  positions their elastic Hessian is the Hessian of the region energy in q.
Pinned by the FD checks of tests/test_oracle.py and tests/test_oracle_adjoint.py (each tangent
against central differences of its gradient) and by central differences of whole oracle
rollouts.
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
        self.ipc = getattr(solver, "ipc", None)
        self.affine = getattr(solver, "affine", None)
        # element moduli of the operator: the solver's, with the rigid regions' own elements
        # (zero in the solver's nodal passes) carrying their moduli, so that the elastic
        # Hessian of a region's elements at the slaved positions is that of its energy in q
        self.mu_el = np.asarray(solver.mu, np.float64).copy()
        self.lam_el = np.asarray(solver.lam, np.float64).copy()
        if self.affine is not None:
            for el in self.affine.elems:
                self.mu_el[el] = self.mu0[el]
                self.lam_el[el] = self.lam0[el]
            self.slaved = np.zeros(self.N, bool)
            self.slaved[self.affine.node] = True

    # reference implicit/adjoint.py:41-50
    def set_density(self, rho):
        E = self.Emin + np.asarray(rho, np.float64) * (1.0 - self.Emin)
        self.mu_el = E * self.mu0
        self.lam_el = E * self.lam0
        self.s.mu = self.mu_el.copy()
        self.s.lam = self.lam_el.copy()
        if self.affine is not None:                  # regions: summed moduli, own elements zero
            self.affine.set_moduli(self.s, self.mu_el, self.lam_el)

    def _cols(self, use_disc, use_sdf):
        return [c for c in self.s.colliders if (use_disc and c["kind"] == 0) or (use_sdf and c["kind"] != 0)]

    # reference implicit/adjoint.py:231-256
    def apply_H(self, x, v, use_disc=False, use_sdf=False, include_mass=True, use_self=False,
                use_ipc=None):
        """H v on node vectors (N, 2). use_ipc (default: on when the solver has IPC contact)
        adds the IPC barrier's tangent between bodies (IPCContact.hvp)."""
        y = np.zeros((self.N, 2))
        K.elastic_hvp(x, v, self.tri, self.B, self.area, self.mu_el, self.lam_el, self.h2, y)
        for c in self._cols(use_disc, use_sdf):
            K.collider_hvp(c, x, v, self.s.bnd, self.s.dhat, self.s.kappa, self.h2, y)
        if use_self and self.s.node is not None:
            self.s.node.hvp(x, v, self.h2, y)
        if self._ipc_on(use_ipc):
            self.ipc.hvp(x, v, self.h2, y)
        if self.pressure is not None:
            self.pressure.hvp(x, v, self.h2, y)
        return self.mass[:, None] * v + y if include_mass else y

    def _ipc_on(self, use_ipc):
        return self.ipc is not None and (use_ipc is None or bool(use_ipc))

    # reference implicit/adjoint.py:356-372
    def diag_precond(self, x, use_disc=False, use_sdf=False, use_self=False, use_ipc=None):
        g = np.zeros((self.N, 2))
        d = np.repeat(self.mass[:, None], 2, axis=1).copy()
        K.elastic_grad_diag(x, self.tri, self.B, self.area, self.mu_el, self.lam_el, self.h2, 1, g, d)
        for c in self._cols(use_disc, use_sdf):
            K.collider_grad_diag(c, x, self.s.bnd, self.s.dhat, self.s.kappa, self.h2, 1, g, d)
        if use_self and self.s.node is not None:
            self.s.node.grad_diag(x, self.h2, 1, g, d)
        if self._ipc_on(use_ipc):
            self.ipc.barrier_grad_diag(self.s, x, self.h2, g, d)
        return np.maximum(d.reshape(-1), 1e-30)

    # ---- reduced coordinates with affine bodies: (free soft degrees of freedom, q per group)
    def _split(self):
        """Mask of the node degrees of freedom kept as they are (free and not slaved)."""
        ff = self.free.reshape(-1) > 0.5
        if self.affine is not None:
            ff &= ~np.repeat(self.slaved, 2)
        return ff

    def expand(self, r):
        """Node vector (N, 2) of a reduced vector r = [soft free values, q (n_groups x 6)]."""
        ff = self._split()
        full = np.zeros((self.N, 2))
        nf = int(ff.sum())
        full.reshape(-1)[ff] = r[:nf]
        if self.affine is not None:
            self.affine.expand_into(r[nf:].reshape(-1, 6), full)
        return full

    def reduce(self, y, square=False):
        """P^T y as a reduced vector (square: the diagonal of P^T diag(y) P)."""
        ff = self._split()
        parts = [y.reshape(-1)[ff]]
        if self.affine is not None:
            parts.append(self.affine.reduce_vec(y, square).reshape(-1))
        return np.concatenate(parts)

    # reference implicit/adjoint.py:375-404
    def adjoint_solve(self, x, rhs, use_disc=False, use_sdf=False, rtol=1e-12, maxiter=20000,
                      use_self=False, use_ipc=None):
        """P^T H P lam_r = P^T rhs by scipy's MINRES with the Jacobi preconditioner; returns the
        node vector P lam_r (N, 2). Without affine bodies P selects the free degrees of freedom
        (the reference's solve)."""
        from scipy.sparse.linalg import LinearOperator, minres
        kw = dict(use_disc=use_disc, use_sdf=use_sdf, use_self=use_self, use_ipc=use_ipc)
        dinv = 1.0 / np.maximum(self.reduce(self.diag_precond(x, **kw).reshape(self.N, 2), square=True),
                                1e-30)
        nr = dinv.size

        def mv(r):
            return self.reduce(self.apply_H(x, self.expand(r), **kw))
        A = LinearOperator((nr, nr), matvec=mv)
        M = LinearOperator((nr, nr), matvec=lambda v: dinv * v)
        lr, info = minres(A, self.reduce(np.asarray(rhs, np.float64).reshape(self.N, 2)), M=M, rtol=rtol,
                          maxiter=maxiter)
        self.info = info
        return self.expand(lr)

    # reference implicit/adjoint.py:165-178
    def drho(self, x, lam):
        F = K.ds_from_nodes(x, self.tri) @ self.B
        HH = self.area[:, None, None] * (K.nh_P(F, self.mu0, self.lam0) @ np.swapaxes(self.B, -1, -2))
        g1, g2 = HH[:, :, 0], HH[:, :, 1]
        l0, l1, l2 = lam[self.tri[:, 0]], lam[self.tri[:, 1]], lam[self.tri[:, 2]]
        s = (l0 * -(g1 + g2)).sum(1) + (l1 * g1).sum(1) + (l2 * g2).sum(1)
        return -(1.0 - self.Emin) * self.h2 * s

    # reference implicit/adjoint.py:373-399
    def transient_adjoint(self, xs, seed, drag, use_disc=False, use_sdf=False, use_self=False, use_ipc=None,
                          **kw):
        T = len(xs) - 1
        m2 = self.mass[:, None]
        lam = {T + 1: np.zeros((self.N, 2)), T + 2: np.zeros((self.N, 2))}
        g = np.zeros(self.n_tri)
        self.dJdp_steps = None if self.pressure is None else np.zeros((T, len(self.pressure.p)))
        for k in range(T, 0, -1):
            rhs = (1.0 + drag) * m2 * lam[k + 1] - drag * m2 * lam[k + 2]
            if k in seed:
                rhs = rhs + seed[k]
            lam[k] = self.adjoint_solve(xs[k], rhs, use_disc, use_sdf, use_self=use_self, use_ipc=use_ipc, **kw)
            g += self.drho(xs[k], lam[k])
            if self.pressure is not None:
                self.dJdp_steps[k - 1] = self.pressure.dp(xs[k], lam[k], self.h2)
        return g
