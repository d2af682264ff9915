"""The "node" contact between bodies in the oracle: a penalty between boundary nodes of one
environment and Coulomb friction on the same pairs (reference implicit/kernels.py:736-869,
implicit/body_contact.py:37-210, adjoint tangent implicit/adjoint.py:115-143).

Broadphase by exact enumeration: for each boundary node b of environment e, every other
boundary node jb of e that is not excluded against it (excl_flat[off_e + i n_be + jl] == 0,
i and jl the two nodes' positions in e's boundary list) and lies at
d = sqrt(|x_b - x_jb|^2 + 1e-12) < r. That is the set the reference's hash-grid query plus its
distance test selects. Every pair is visited from both of its ends. A visit adds its own node's
gradient and diagonal, and half of the pair's p^T H p and energy:

    E     = 1/2 k (r - d)^2                     per pair (1/4 per visit)
    g_b   = -k (r - d) n,   n = (x_b - x_jb) / d
    H_bb  = k n n^T - k (r - d) / d (I - n n^T)  (diagonal entries clamped at 0 in the solve)

Friction on a pair uses the relative slip over the step u = (x_b - xn_b) - (x_jb - xn_jb),
f_n = k (r - d) and the smoothed Coulomb law of the colliders (reference
implicit/body_contact.py:1-27).

Test infrastructure only: the product never imports this.
"""
from __future__ import annotations

import numpy as np

REG = 1.0e-12                      # d = sqrt(|diff|^2 + REG)  (reference implicit/kernels.py:793)


def pair_mu(mu, mu_node, i, j):
    """Coefficient of a contact between the bodies of nodes i and j (reference
    core/math.py:133-144): mu, or with per-body values the harmonic mean 2 m_i m_j / (m_i + m_j),
    0 when both are 0."""
    if mu_node is None:
        return np.full(np.shape(i), float(mu))
    mi, mj = mu_node[i], mu_node[j]
    sm = mi + mj
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(sm > 0.0, 2.0 * mi * mj / np.where(sm > 0.0, sm, 1.0), 0.0)


class NodeContact:
    """Penalty contact between the boundary nodes of one environment (reference
    implicit/kernels.py:755-869, launched from implicit/solver.py:279-285, 370-375, 475-481,
    591-597)."""
    in_adjoint = True

    def __init__(self, s, stiffness, radius):
        me = s.me
        self.k, self.r = float(stiffness), float(radius)
        self.bnd = np.asarray(me.boundary_nodes, np.int64)
        self.b2e = np.asarray(me.boundary2env, np.int64)
        self.bstart = np.asarray(me.boundary_starts, np.int64)
        self.bpe = np.asarray(me.boundary_per_env, np.int64)
        self.excl_flat = np.asarray(me.excl_flat, np.uint8)
        self.excl_off = np.asarray(me.excl_offset, np.int64)
        self.B = s.B

    # ---- broadphase (every visit, ordered by visiting node)
    def visits(self, x):
        """(node, other, n, d) of every visit at positions x: node and other are node ids, n the
        unit vector from other to node, d the regularised distance."""
        vb, vj = [], []
        for e in range(self.B):
            st, n = int(self.bstart[e]), int(self.bpe[e])
            if n < 2:
                continue
            P = x[self.bnd[st:st + n]]
            D = P[:, None, :] - P[None, :, :]
            dd = np.sqrt((D * D).sum(-1) + REG)
            ex = self.excl_flat[self.excl_off[e]:self.excl_off[e] + n * n].reshape(n, n) != 0
            ok = (dd < self.r) & ~ex
            np.fill_diagonal(ok, False)
            bi, bj = np.nonzero(ok)
            vb.append(st + bi)
            vj.append(st + bj)
        b = np.concatenate(vb) if vb else np.zeros(0, np.int64)
        j = np.concatenate(vj) if vj else np.zeros(0, np.int64)
        node, other = self.bnd[b], self.bnd[j]
        diff = x[node] - x[other]
        d = np.sqrt((diff * diff).sum(1) + REG)
        return node, other, diff / d[:, None], d, self.b2e[b]

    # ---- grad_diag_self_bp_kernel (reference implicit/kernels.py:772-808)
    def grad_diag(self, x, h2, clamp, grad, diag):
        node, _, nh, d, _ = self.visits(x)
        rd = self.r - d
        np.add.at(grad, node, h2 * (-self.k * rd)[:, None] * nh)
        tang = self.k * rd / d
        hx = self.k * nh[:, 0] ** 2 - tang * (1.0 - nh[:, 0] ** 2)
        hy = self.k * nh[:, 1] ** 2 - tang * (1.0 - nh[:, 1] ** 2)
        if clamp:
            hx, hy = np.maximum(hx, 0.0), np.maximum(hy, 0.0)
        np.add.at(diag, node, h2 * np.stack([hx, hy], 1))

    # ---- pHp_self_bp_kernel (reference implicit/kernels.py:811-845)
    def pHp(self, x, p, h2, clamp, pHp):
        node, other, nh, d, env = self.visits(x)
        rd = self.r - d
        dp = p[node] - p[other]
        pn = (dp * nh).sum(1)
        pt = dp - pn[:, None] * nh
        v = self.k * pn * pn - self.k * rd / d * (pt * pt).sum(1)
        if clamp:
            v = np.maximum(v, 0.0)
        np.add.at(pHp, env, h2 * 0.5 * v)

    # ---- energy_self_bp_kernel (reference implicit/kernels.py:848-869)
    def energy(self, x, h2, E):
        _, _, _, d, env = self.visits(x)
        rd = self.r - d
        np.add.at(E, env, h2 * 0.5 * (0.5 * self.k * rd * rd))

    # ---- self_bp_hvp_kernel (reference implicit/adjoint.py:115-143): own rows, unclamped
    def hvp(self, x, v, h2, y):
        node, other, nh, d, _ = self.visits(x)
        rd = self.r - d
        dv = v[node] - v[other]
        vn = (dv * nh).sum(1)
        vt = dv - vn[:, None] * nh
        np.add.at(y, node, h2 * (self.k * vn[:, None] * nh - (self.k * rd / d)[:, None] * vt))


class BodyFriction:
    """Coulomb friction on the node-contact pairs (reference implicit/body_contact.py:170-210)."""
    in_adjoint = False
    persistent = True

    def __init__(self, s, contact, mu, mu_node=None):
        self.c = contact
        self.mu = float(mu)
        self.mu_node = None if mu_node is None else np.asarray(mu_node, np.float64)

    def _frame(self, s, x):
        node, other, nh, d, env = self.c.visits(x)
        f_n = self.c.k * (self.c.r - d)
        u = (x[node] - s.xn[node]) - (x[other] - s.xn[other])
        return node, other, nh, f_n, u, env, pair_mu(self.mu, self.mu_node, node, other)

    # reference implicit/body_contact.py:37-82
    def _grad_diag(self, s, h2, free, grad, diag):
        node, _, nh, f_n, u, _, mu = self._frame(s, s.x)
        v = u * s.inv_dt
        vt = v - (v * nh).sum(1)[:, None] * nh
        den = np.maximum(np.linalg.norm(vt, axis=1), s.epsv)
        g = (mu * f_n / den)[:, None] * vt
        kt = mu * f_n * s.inv_dt / den
        dg = np.stack([kt * nh[:, 1] ** 2, kt * nh[:, 0] ** 2], 1)
        np.add.at(grad, node, h2 * g * free[node])
        np.add.at(diag, node, h2 * dg * free[node])

    def grad_diag(self, s):
        self._grad_diag(s, s.h2, s.free, s.grad, s.diagH)

    def pair_force(self, s, grad, diag, free):
        self._grad_diag(s, 1.0, free, grad, diag)

    # reference implicit/body_contact.py:86-124
    def pHp(self, s):
        node, other, nh, f_n, u, env, mu = self._frame(s, s.x)
        v = u * s.inv_dt
        vt = v - (v * nh).sum(1)[:, None] * nh
        kt = mu * f_n * s.inv_dt / np.maximum(np.linalg.norm(vt, axis=1), s.epsv)
        dp = s.p[node] - s.p[other]
        pt = dp[:, 1] * nh[:, 0] - dp[:, 0] * nh[:, 1]
        np.add.at(s.pHp, env, s.h2 * 0.5 * kt * pt * pt)

    # reference implicit/body_contact.py:128-167
    def energy(self, s, buf):
        _, _, nh, f_n, u, env, mu = self._frame(s, s.x)
        a = s.epsv / s.inv_dt
        ut = np.abs(u[:, 1] * nh[:, 0] - u[:, 0] * nh[:, 1])
        f0 = np.where(ut < a, ut * ut / (2.0 * a), ut - 0.5 * a)
        np.add.at(buf, env, s.h2 * 0.5 * mu * f_n * f0)
