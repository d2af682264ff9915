"""Body-level modes and actuators of the oracle: affine (rigid) bodies, the rigid-mode coarse
space, external force, pressure and tendons (reference implicit/affine.py,
implicit/actuation.py). Test infrastructure only.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from .terms import nh_P, nh_d2, det2, inv2, ddot, LOG_EPS

INV_SQRT2 = 0.7071067811865476
_AIDX = (1, 2, 4, 5)             # q entries holding A11, A12, A21, A22


def A_of(q):
    return np.stack([np.stack([q[:, 1], q[:, 2]], -1), np.stack([q[:, 4], q[:, 5]], -1)], 1)


def reduced_hessian(A, Mu, La, h2, Hx, Hy):
    """J^T D J + h2 d2Psi/dA2 with the negative twist direction removed, per group
    (reference implicit/affine.py:61-101)."""
    n = A.shape[0]
    H = np.zeros((n, 6, 6))
    H[:, 0:3, 0:3] = Hx
    H[:, 3:6, 3:6] = Hy
    for k in range(4):
        E = np.zeros((n, 2, 2))
        E[:, k // 2, k % 2] = 1.0
        col = nh_d2(A, E, Mu, La)
        cv = col.reshape(n, 4)
        for r in range(4):
            H[:, _AIDX[r], _AIDX[k]] += h2 * cv[:, r]
    a = A[:, 0, 0] + A[:, 1, 1]
    b = A[:, 1, 0] - A[:, 0, 1]
    nn = np.sqrt(a * a + b * b)
    ok = nn > 1.0e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.stack([-b, -a, a, -b], 1) * (INV_SQRT2 / nn)[:, None]
    Tm = t.reshape(n, 2, 2)
    lam_t = ddot(Tm, nh_d2(A, Tm, Mu, La))
    fix = ok & (lam_t < 0.0)
    for r in range(4):
        for c in range(4):
            H[fix, _AIDX[r], _AIDX[c]] -= h2 * lam_t[fix] * t[fix, r] * t[fix, c]
    return H


def chol6(H):
    """Lower Cholesky factors; a group that is not positive definite gets the factor of
    its diagonal (reference implicit/affine.py:105-129)."""
    n = H.shape[0]
    L = np.zeros((n, 6, 6))
    ok = np.ones(n, bool)
    for j in range(6):
        s = H[:, j, j] - (L[:, j, :j] ** 2).sum(1)
        ok &= ~((s <= 1.0e-9 * np.abs(H[:, j, j])) | (s <= 0.0))
        d = np.sqrt(np.where(ok, s, 1.0))
        L[:, j, j] = np.where(ok, d, L[:, j, j])
        for i in range(j + 1, 6):
            v = H[:, i, j] - (L[:, i, :j] * L[:, j, :j]).sum(1)
            L[:, i, j] = np.where(ok, v / d, L[:, i, j])
    bad = ~ok
    if bad.any():
        L[bad] = 0.0
        for j in range(6):
            L[bad, j, j] = np.sqrt(np.maximum(H[bad, j, j], 1.0e-30))
    return L


def chol6_solve(L, b):
    n = L.shape[0]
    y = np.zeros((n, 6))
    for i in range(6):
        y[:, i] = (b[:, i] - (L[:, i, :i] * y[:, :i]).sum(1)) / L[:, i, i]
    x = np.zeros((n, 6))
    for i in range(5, -1, -1):
        x[:, i] = (y[:, i] - (L[:, i + 1:, i] * x[:, i + 1:]).sum(1)) / L[:, i, i]
    return x


class AffineBodies:
    """Rigid regions in reduced coordinates q = (t_x, A11, A12, t_y, A21, A22)
    (reference implicit/affine.py:272-400)."""
    in_adjoint = False
    persistent = True

    def __init__(self, s, groups):
        me = s.me
        tri = np.asarray(me.tri)
        mass = np.asarray(me.mass, np.float64)
        X = np.asarray(me.rest_pos, np.float64)
        area = np.asarray(me.rest_area, np.float64)
        node, grp, w, Mu, La, env = [], [], [], [], [], []
        owner = np.full(X.shape[0], -1)
        self.elems = [np.asarray(g["elems"], np.int64) for g in groups]
        self.area = area
        for b, g in enumerate(groups):
            el = np.asarray(g["elems"], np.int64)
            nd = np.unique(tri[el].reshape(-1))
            if (owner[nd] >= 0).any():
                raise ValueError("two rigid regions share a node")
            if (s.free[nd] == 0).any():
                raise ValueError("a rigid region contains kinematically constrained nodes")
            owner[nd] = b
            c = (mass[nd, None] * X[nd]).sum(0) / mass[nd].sum()
            node += list(nd)
            grp += [b] * len(nd)
            w.append(np.c_[np.ones(len(nd)), X[nd] - c])
            Mu.append(float((area[el] * np.asarray(me.mu, np.float64)[el]).sum()))
            La.append(float((area[el] * np.asarray(me.lam, np.float64)[el]).sum()))
            env.append(int(np.asarray(me.tri2env)[el[0]]))
            s.mu[el] = 0.0
            s.lam[el] = 0.0
        self.n = len(groups)
        self.node = np.asarray(node, np.int64)
        self.grp = np.asarray(grp, np.int64)
        self.w = np.concatenate(w, 0)
        self.first = np.concatenate([[0], np.cumsum([len(a) for a in w])])
        self.Mu, self.La = np.asarray(Mu), np.asarray(La)
        self.env = np.asarray(env, np.int64)
        z = lambda: np.zeros((self.n, 6))
        self.q, self.q0, self.pq, self.gq, self.gq_prev, self.zq = z(), z(), z(), z(), z(), z()
        self.Hx = np.zeros((self.n, 3, 3))
        self.Hy = np.zeros((self.n, 3, 3))
        self.slaved = owner >= 0
        self.sync_from_x(s)

    def sync_from_x(self, s, regions=None):
        q = np.zeros((self.n, 6)) if regions is None else self.q.copy()
        for b in (range(self.n) if regions is None else regions):
            k = slice(self.first[b], self.first[b + 1])
            sol = np.linalg.lstsq(self.w[k], s.x[self.node[k]], rcond=None)[0]
            q[b, :3], q[b, 3:] = sol[:, 0], sol[:, 1]
        self.q = q
        self.pq[:] = 0.0
        self.gq[:] = 0.0
        self.gq_prev[:] = 0.0
        self._expand(self.q, s.x)

    def _expand(self, v6, out):
        v = v6[self.grp]
        out[self.node, 0] = (self.w * v[:, 0:3]).sum(1)
        out[self.node, 1] = (self.w * v[:, 3:6]).sum(1)

    def begin_step(self, s):
        pass

    def after_elastic(self, s):
        pass

    # ---- the adjoint's reduction (no counterpart in the reference, whose adjoint refuses
    # affine bodies). The slaved nodes follow x_i = J_i q, so the implicit step's optimality
    # conditions are those of the energy in (free soft nodes, q): its Hessian is P^T H P with
    # P the map from those coordinates to node positions.
    def set_moduli(self, s, mu_full, lam_full):
        """Region energies from per-element moduli (the area-weighted sums); the region's own
        elements stay zero in the nodal passes."""
        for b, el in enumerate(self.elems):
            self.Mu[b] = float((self.area[el] * mu_full[el]).sum())
            self.La[b] = float((self.area[el] * lam_full[el]).sum())
            s.mu[el] = 0.0
            s.lam[el] = 0.0

    def expand_into(self, q6, out):
        """out[slaved nodes] = J q (q: (n, 6))."""
        self._expand(q6, out)
        return out

    def reduce_vec(self, y, square=False):
        """P^T y per group ((n, 6)): sum over the group's nodes of w y_x, w y_y; with square,
        the diagonal of J^T diag(y) J (weights squared)."""
        w = self.w * self.w if square else self.w
        yn = y[self.node]
        out = np.zeros((self.n, 6))
        np.add.at(out, self.grp, np.concatenate([yn[:, 0:1] * w, yn[:, 1:2] * w], 1))
        return out

    def gather(self, s):
        self.gq_prev = self.gq.copy()
        self.gq[:] = 0.0
        self.Hx[:] = 0.0
        self.Hy[:] = 0.0
        g = s.grad[self.node]
        d = s.diagH[self.node]
        gq = np.concatenate([g[:, 0:1] * self.w, g[:, 1:2] * self.w], 1)
        np.add.at(self.gq, self.grp, gq)
        o = self.w[:, :, None] * self.w[:, None, :]
        np.add.at(self.Hx, self.grp, d[:, 0, None, None] * o)
        np.add.at(self.Hy, self.grp, d[:, 1, None, None] * o)
        s.grad[self.node] = 0.0

    def reduce(self, s):
        A = A_of(self.q)
        P = nh_P(A, self.Mu, self.La)
        g = self.gq + s.h2 * np.stack([np.zeros(self.n), P[:, 0, 0], P[:, 0, 1],
                                       np.zeros(self.n), P[:, 1, 0], P[:, 1, 1]], 1)
        L = chol6(reduced_hessian(A, self.Mu, self.La, s.h2, self.Hx, self.Hy))
        y = g - self.gq_prev
        z = chol6_solve(L, g)
        zy = chol6_solve(L, y)
        pp = self.pq
        np.add.at(s.gPg, self.env, (g * z).sum(1))
        np.add.at(s.g_p, self.env, (g * pp).sum(1))
        np.add.at(s.g_Py, self.env, (g * zy).sum(1))
        np.add.at(s.y_p, self.env, (y * pp).sum(1))
        np.add.at(s.y_Py, self.env, (y * zy).sum(1))
        self.gq = g
        self.zq = z

    def direction(self, s):
        self.pq = -self.zq + s.beta[self.env][:, None] * self.pq
        self._expand(self.pq, s.p)

    def advance(self, s):
        act = s.active[self.env] != 0
        self.q[act] = self.q[act] + s.alpha[self.env[act]][:, None] * self.pq[act]
        self._expand(self.q, s.x)

    def save(self, s):
        self.q0 = self.q.copy()

    def trial(self, s):
        act = s.active[self.env] != 0
        self.q[act] = self.q0[act] + s.alpha[self.env[act]][:, None] * self.pq[act]
        self._expand(self.q, s.x)

    def grad_diag(self, s):
        pass

    def pHp(self, s):
        dA = A_of(self.pq)
        val = ddot(dA, nh_d2(A_of(self.q), dA, self.Mu, self.La))
        np.add.at(s.pHp, self.env, s.h2 * np.maximum(val, 0.0))

    def energy(self, s, buf):
        A = A_of(self.q)
        lj = np.log(np.maximum(det2(A), LOG_EPS))
        psi = 0.5 * self.Mu * (ddot(A, A) - 2.0) - self.Mu * lj + 0.5 * self.La * lj * lj
        np.add.at(buf, self.env, s.h2 * psi)

    def state(self):
        return self.q[:, [0, 3]], np.stack([self.q[:, 1:3], self.q[:, 4:6]], 1)


class RigidModes:
    """Translation + rotation coarse space of every free soft piece
    (reference implicit/affine.py:489-563)."""

    @classmethod
    def build(cls, s, skip_nodes=None):
        tri = np.asarray(s.me.tri)
        a = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
        b = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
        ncomp, lab = connected_components(coo_matrix((np.ones(len(a)), (a, b)), shape=(s.N, s.N)),
                                          directed=False)
        held = np.zeros(ncomp, bool)
        held[np.unique(lab[(s.free == 0).any(1)])] = True
        if skip_nodes is not None and np.any(skip_nodes):
            held[np.unique(lab[np.asarray(skip_nodes, bool)])] = True
        size = np.bincount(lab, minlength=ncomp)
        keep = np.where(~held & (size >= 3))[0]
        if len(keep) == 0:
            return None
        newid = np.full(ncomp, -1)
        newid[keep] = np.arange(len(keep))
        nodes = np.where(newid[lab] >= 0)[0]
        return cls(s, nodes, newid[lab[nodes]], len(keep))

    def __init__(self, s, nodes, grp, n):
        self.n = int(n)
        self.node, self.grp = np.asarray(nodes), np.asarray(grp)
        first = np.zeros(n, np.int64)
        first[self.grp[::-1]] = self.node[::-1]
        self.env = np.asarray(s.node2env)[first]
        self.diag_el = np.zeros((s.N, 2))

    def begin_step(self, s):
        pass

    def after_elastic(self, s):
        self.diag_el = s.diagH.copy()

    def _arm(self, s):
        c = self.acc[:, :2] / self.acc[:, 2:3]
        return s.x[self.node] - c[self.grp]

    def gather(self, s):
        self.acc = np.zeros((self.n, 3))
        m = s.mass[self.node]
        xi = s.x[self.node]
        np.add.at(self.acc, self.grp, np.stack([m * xi[:, 0], m * xi[:, 1], m], 1))
        rr = self._arm(s)
        g = s.grad[self.node]
        gp = s.grad_prev[self.node]
        d = s.diagH[self.node] - self.diag_el[self.node]
        dx = np.maximum(d[:, 0], 0.0) + m
        dy = np.maximum(d[:, 1], 0.0) + m
        self.gq = np.zeros((self.n, 3))
        self.gq_prev = np.zeros((self.n, 3))
        self.Hc = np.zeros((self.n, 3, 3))
        np.add.at(self.gq, self.grp, np.stack([g[:, 0], g[:, 1], rr[:, 0] * g[:, 1] - rr[:, 1] * g[:, 0]], 1))
        np.add.at(self.gq_prev, self.grp,
                  np.stack([gp[:, 0], gp[:, 1], rr[:, 0] * gp[:, 1] - rr[:, 1] * gp[:, 0]], 1))
        Hc = np.zeros((len(self.node), 3, 3))
        Hc[:, 0, 0] = dx
        Hc[:, 0, 2] = -dx * rr[:, 1]
        Hc[:, 1, 1] = dy
        Hc[:, 1, 2] = dy * rr[:, 0]
        Hc[:, 2, 0] = -dx * rr[:, 1]
        Hc[:, 2, 1] = dy * rr[:, 0]
        Hc[:, 2, 2] = dx * rr[:, 1] ** 2 + dy * rr[:, 0] ** 2
        np.add.at(self.Hc, self.grp, Hc)

    def reduce(self, s):
        Hi = np.linalg.inv(self.Hc)
        g = self.gq
        y = g - self.gq_prev
        z = np.einsum("nij,nj->ni", Hi, g)
        zy = np.einsum("nij,nj->ni", Hi, y)
        np.add.at(s.gPg, self.env, (g * z).sum(1))
        np.add.at(s.g_Py, self.env, (g * zy).sum(1))
        np.add.at(s.y_Py, self.env, (y * zy).sum(1))
        self.zq = z

    def direction(self, s):
        z = self.zq[self.grp]
        rr = self._arm(s)
        s.p[self.node, 0] -= z[:, 0] - z[:, 2] * rr[:, 1]
        s.p[self.node, 1] -= z[:, 1] + z[:, 2] * rr[:, 0]

    def advance(self, s):
        pass

    def save(self, s):
        pass

    def trial(self, s):
        pass


class ExternalForce:
    """Nodal force field f: grad -= h2 f, energy -h2 f.(x - xn) (reference implicit/actuation.py:37-72)."""
    in_adjoint = True

    def __init__(self, s):
        self.f = np.zeros((s.N, 2))

    def set(self, f):
        self.f = np.asarray(f, np.float64).copy()

    def grad_diag(self, s):
        s.grad -= s.h2 * self.f

    def pHp(self, s):
        pass

    def energy(self, s, buf):
        np.add.at(buf, s.node2env, -s.h2 * (self.f * (s.x - s.xn)).sum(1))


class Pressure:
    """E = -p A(x) per chamber (reference implicit/actuation.py:89-151)."""
    in_adjoint = False

    def __init__(self, s, ea, eb, ech, ch2env, cref):
        self.ea, self.eb, self.ech = (np.asarray(v, np.int64) for v in (ea, eb, ech))
        self.ch2env = np.asarray(ch2env, np.int64)
        self.cref = np.asarray(cref, np.float64)
        self.p = np.zeros(len(self.ch2env))

    def set(self, p):
        self.p = np.asarray(p, np.float64).copy()

    def grad_diag(self, s):
        t = s.x[self.eb] - s.x[self.ea]
        n = np.stack([t[:, 1], -t[:, 0]], 1)
        g = (-0.5 * s.h2 * self.p[self.ech])[:, None] * n
        np.add.at(s.grad, self.ea, g)
        np.add.at(s.grad, self.eb, g)

    def pHp(self, s):
        pa, pb = s.p[self.ea], s.p[self.eb]
        c = self.ech
        np.add.at(s.pHp, self.ch2env[c], -s.h2 * self.p[c] * (pa[:, 0] * pb[:, 1] - pa[:, 1] * pb[:, 0]))

    def energy(self, s, buf):
        c = self.ech
        ra = s.x[self.ea] - self.cref[c]
        rb = s.x[self.eb] - self.cref[c]
        np.add.at(buf, self.ch2env[c], -s.h2 * self.p[c] * 0.5 * (ra[:, 0] * rb[:, 1] - ra[:, 1] * rb[:, 0]))

    # The adjoint's pieces. Not in the reference, whose adjoint refuses Pressure
    # (implicit/adjoint.py:376-385); derived from its gradient above: each edge adds
    # -1/2 h2 p J(x_b - x_a) to both ends, J(u) = (u_y, -u_x).
    def hvp(self, x, v, h2, y):
        """y += h2 H_pressure v: the directional derivative of the gradient along v (linear in
        x, so exact), w = -1/2 h2 p J(v_b - v_a) on both ends of every edge."""
        u = v[self.eb] - v[self.ea]
        w = (-0.5 * h2 * self.p[self.ech])[:, None] * np.stack([u[:, 1], -u[:, 0]], 1)
        np.add.at(y, self.ea, w)
        np.add.at(y, self.eb, w)

    def dp(self, x, lam, h2):
        """-lam . dR/dp per chamber, R the gradient of the incremental potential: dR/dp_c puts
        -1/2 h2 J(x_b - x_a) on both ends of chamber c's edges, so this is
        1/2 h2 sum_e (lam_a + lam_b) . J(x_b - x_a)."""
        t = x[self.eb] - x[self.ea]
        Jt = np.stack([t[:, 1], -t[:, 0]], 1)
        out = np.zeros(len(self.ch2env))
        np.add.at(out, self.ech, 0.5 * h2 * ((lam[self.ea] + lam[self.eb]) * Jt).sum(1))
        return out


def _seg_dir(xa, xb):
    t = xb - xa
    ln = np.maximum(np.linalg.norm(t, axis=1), 1.0e-12)
    return t / ln[:, None], 1.0 / ln


class Tendon:
    """Cables through eyelets: E = T L (tension) or k/2 (L - L0)+^2 (length)
    (reference implicit/actuation.py:169-316)."""
    in_adjoint = False

    def __init__(self, s, node, prv, nxt, tid, t2env, stiff, L0):
        self.node, self.prv, self.nxt, self.tid = (np.asarray(v, np.int64) for v in (node, prv, nxt, tid))
        self.t2env = np.asarray(t2env, np.int64)
        self.stiff = np.asarray(stiff, np.float64)
        self.L0 = np.asarray(L0, np.float64)
        self.Tcmd = np.zeros(len(self.t2env))
        self.tau = np.zeros(len(self.t2env))
        self.keff = np.zeros(len(self.t2env))

    def set(self, tension, length):
        self.Tcmd = np.asarray(tension, np.float64).copy()
        self.L0 = np.asarray(length, np.float64).copy()

    def _lengths(self, s):
        L = np.zeros(len(self.t2env))
        j = np.nonzero(self.nxt >= 0)[0]
        seg = np.linalg.norm(s.x[self.node[self.nxt[j]]] - s.x[self.node[j]], axis=1)
        np.add.at(L, self.tid[j], seg)
        return L

    def grad_diag(self, s):
        L = self._lengths(s)
        taut = self.stiff > 0.0
        ext = L - self.L0
        self.tau = np.where(taut, np.where(ext > 0.0, self.stiff * ext, 0.0), np.maximum(self.Tcmd, 0.0))
        self.keff = np.where(taut & (ext > 0.0), self.stiff, 0.0)
        n = self.node
        gL = np.zeros((len(n), 2))
        geo = np.zeros((len(n), 2))
        hp = self.prv >= 0
        d, il = _seg_dir(s.x[self.node[self.prv[hp]]], s.x[n[hp]])
        gL[hp] += d
        geo[hp] += (1.0 - d * d) * il[:, None]
        hn = self.nxt >= 0
        d, il = _seg_dir(s.x[n[hn]], s.x[self.node[self.nxt[hn]]])
        gL[hn] -= d
        geo[hn] += (1.0 - d * d) * il[:, None]
        k = self.tid
        np.add.at(s.grad, n, (s.h2 * self.tau[k])[:, None] * gL)
        np.add.at(s.diagH, n, s.h2 * (self.tau[k][:, None] * geo + self.keff[k][:, None] * gL * gL))

    def pHp(self, s):
        j = np.nonzero(self.nxt >= 0)[0]
        k = self.tid[j]
        d, il = _seg_dir(s.x[self.node[j]], s.x[self.node[self.nxt[j]]])
        dp = s.p[self.node[self.nxt[j]]] - s.p[self.node[j]]
        along = (dp * d).sum(1)
        dL = np.zeros(len(self.t2env))
        np.add.at(dL, k, along)
        np.add.at(s.pHp, self.t2env[k], s.h2 * self.tau[k] * ((dp * dp).sum(1) - along * along) * il)
        np.add.at(s.pHp, self.t2env, s.h2 * self.keff * dL * dL)

    def energy(self, s, buf):
        L = self._lengths(s)
        ext = L - self.L0
        e = np.where(self.stiff > 0.0, np.where(ext > 0.0, 0.5 * s.h2 * self.stiff * ext * ext, 0.0),
                     s.h2 * np.maximum(self.Tcmd, 0.0) * L)
        np.add.at(buf, self.t2env, e)


class Springs:
    """Linear springs and pins, E = k/2 (|d| - L0)^2 with d = x_a - x_b, x_b a world anchor
    when b = -1 (reference implicit/actuation.py:368-476). The tangential stiffness
    k (1 - L0/|d|) is clamped at 0 in the diagonal and in p^T H p."""
    in_adjoint = False

    def __init__(self, s, node_a, node_b, stiffness, s2env):
        self.a = np.asarray(node_a, np.int64)
        self.b = np.asarray(node_b, np.int64)
        self.k = np.asarray(stiffness, np.float64)
        self.env = np.asarray(s2env, np.int64)
        self.L0 = np.zeros(len(self.a))
        self.anchor = np.zeros((len(self.a), 2))

    def set(self, anchor, length):
        self.anchor = np.asarray(anchor, np.float64).reshape(-1, 2).copy()
        self.L0 = np.asarray(length, np.float64).copy()

    def _geom(self, x):
        """(n, tension, tangential stiffness, length) of every spring (reference :386-392)."""
        world = self.b < 0
        xb = np.where(world[:, None], self.anchor, x[np.maximum(self.b, 0)])
        d = x[self.a] - xb
        ln = np.sqrt((d * d).sum(1) + 1.0e-30)
        return d / ln[:, None], self.k * (ln - self.L0), np.maximum(self.k * (1.0 - self.L0 / ln), 0.0), ln

    def _rel(self, v):
        """v_a - v_b per spring (v_a for a spring to the world)."""
        out = v[self.a].copy()
        both = self.b >= 0
        out[both] -= v[self.b[both]]
        return out

    # reference implicit/actuation.py:395-413
    def grad_diag(self, s):
        n, ten, tang, _ = self._geom(s.x)
        g = s.h2 * ten[:, None] * n
        hd = s.h2 * (self.k[:, None] * n * n + tang[:, None] * (1.0 - n * n))
        np.add.at(s.grad, self.a, g)
        np.add.at(s.diagH, self.a, hd)
        both = self.b >= 0
        np.add.at(s.grad, self.b[both], -g[both])
        np.add.at(s.diagH, self.b[both], hd[both])

    # reference implicit/actuation.py:416-434
    def pHp(self, s):
        n, _, tang, _ = self._geom(s.x)
        dp = self._rel(s.p)
        along = (dp * n).sum(1)
        np.add.at(s.pHp, self.env, s.h2 * (self.k * along * along + tang * ((dp * dp).sum(1) - along * along)))

    # reference implicit/actuation.py:437-446
    def energy(self, s, buf):
        _, _, _, ln = self._geom(s.x)
        ext = ln - self.L0
        np.add.at(buf, self.env, 0.5 * s.h2 * self.k * ext * ext)


class ActiveStrain:
    """Active material: the rest shape of each region's elements becomes A Dm, so the elastic
    term sees Bm = Bm0 A^-1 and area = area0 det A (reference implicit/actuation.py:319-366). No
    energy term of its own."""
    in_adjoint = False

    def __init__(self, s, elem, ereg, n_regions):
        self.elem = np.asarray(elem, np.int64)
        self.ereg = np.asarray(ereg, np.int64)
        self.n_reg = int(n_regions)
        if getattr(s, "Bm0", None) is None:
            s.Bm0, s.area0 = s.Bm.copy(), s.area.copy()

    def set(self, s, A):
        A = np.asarray(A, np.float64).reshape(-1, 2, 2)
        Ainv, detA = np.linalg.inv(A), np.linalg.det(A)
        s.Bm[self.elem] = s.Bm0[self.elem] @ Ainv[self.ereg]
        s.area[self.elem] = s.area0[self.elem] * detA[self.ereg]

    def grad_diag(self, s):
        pass

    def pHp(self, s):
        pass

    def energy(self, s, buf):
        pass
