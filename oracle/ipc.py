"""IPC node-edge contact between bodies in the oracle (reference implicit/ipc_contact.py).

The broadphase is exact enumeration: every (boundary edge, boundary node) pair of one
environment, kept when the node is not an end of the edge, not excluded against either end,
and its distance to the segment is below `reach` — the pair set the reference's hash-grid
query plus its exact-distance filter selects. Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

from .terms import barrier_b, barrier_bp, barrier_bpp, cross2

BC_FLOOR = 0.25


def seg_dist(p, a, b):
    """Distance from p to segment ab (reference implicit/ipc_contact.py:95-99)."""
    ab = b - a
    t = np.clip(((p - a) * ab).sum(-1) / np.maximum((ab * ab).sum(-1), 1.0e-30), 0.0, 1.0)
    return np.linalg.norm(p - a - t[..., None] * ab, axis=-1)


def accd(p, a, b, dp, da, db, g, t_c):
    """Additive CCD of nodes against edges, vectorised over pairs
    (reference implicit/ipc_contact.py:103-127)."""
    m = (dp + da + db) / 3.0
    dp, da, db = dp - m, da - m, db - m
    lp = np.linalg.norm(dp, axis=1) + np.maximum(np.linalg.norm(da, axis=1), np.linalg.norm(db, axis=1))
    out = np.array(t_c, np.float64, copy=True)
    live = lp > 1.0e-30
    d = seg_dist(p, a, b)
    t = np.zeros(len(p))
    with np.errstate(divide="ignore", invalid="ignore"):
        tl = (d - g) / lp
    for k in range(32):
        if not live.any():
            break
        tn = t + tl
        hit_c = live & (tn >= t_c)
        out[hit_c] = t_c[hit_c]
        live &= ~hit_c
        with np.errstate(invalid="ignore", over="ignore"):
            dn = seg_dist(p + tn[:, None] * dp, a + tn[:, None] * da, b + tn[:, None] * db)
        if k > 0:
            stop = live & (dn < g)
            out[stop] = t[stop]
            live &= ~stop
        t = np.where(live, tn, t)
        with np.errstate(divide="ignore", invalid="ignore"):
            tl = np.where(live, 0.9 * dn / lp, tl)
    out[live] = t[live]
    return out


class IPCContact:
    """Node-edge IPC barrier with lagged friction, vertex corrections, ACCD step limit and the
    kinematic sweep (reference implicit/ipc_contact.py:343-483)."""
    in_adjoint = False
    persistent = True

    def __init__(self, s, mu, dhat, kappa, reach, gap=0.1, mu_node=None):
        me = s.me
        edges = np.asarray(me.boundary_edges, np.int64)
        self.ea, self.eb = edges[:, 0], edges[:, 1]
        self.e2env = np.asarray(me.node2env)[self.ea]
        bstart = np.asarray(me.boundary_starts)
        pos = np.full(s.N, -1, np.int64)
        pos[np.asarray(me.boundary_nodes)] = np.arange(s.n_b) - bstart[np.asarray(me.boundary2env)]
        self.ela, self.elb = pos[self.ea], pos[self.eb]
        m = np.bincount(edges.reshape(-1), minlength=s.N).astype(np.float64)
        share = (m - 1.0) / np.maximum(m, 1.0)
        self.ca, self.cb = share[self.ea], share[self.eb]
        self.mu, self.dhat, self.kappa = float(mu), float(dhat), float(kappa)
        self.per_body = mu_node is not None
        self.mu_node = None if mu_node is None else np.asarray(mu_node, np.float64)
        self.reach, self.gap = float(reach), float(gap)
        self.bnd = np.asarray(me.boundary_nodes, np.int64)
        self.bstart = bstart
        self.excl = [np.asarray(e, bool) for e in me.excluded_pairs_per_env]
        self.excl_flat = np.asarray(me.excl_flat)
        self.excl_off = np.asarray(me.excl_offset)
        self.bpe = np.asarray(me.boundary_per_env)
        self.B = s.B
        self.clr = np.zeros(s.B)
        self.q = None
        self.pending = False

    # ---- candidate pairs
    def pairs(self, x):
        """(edge k, boundary index jb, tt, dvec, d) of every pair within reach."""
        ks, jbs = [], []
        for e in range(self.B):
            kk = np.nonzero(self.e2env == e)[0]
            lo, hi = int(self.bstart[e]), int(self.bstart[e + 1])
            if kk.size == 0 or hi == lo:
                continue
            jb = np.arange(lo, hi)
            K, JB = np.meshgrid(kk, jb, indexing="ij")
            K, JB = K.ravel(), JB.ravel()
            pn = self.bnd[JB]
            ok = (pn != self.ea[K]) & (pn != self.eb[K])
            n_be = int(self.bpe[e])
            off = int(self.excl_off[e])
            jl = JB - lo
            ok &= (self.excl_flat[off + jl * n_be + self.ela[K]] == 0)
            ok &= (self.excl_flat[off + jl * n_be + self.elb[K]] == 0)
            ks.append(K[ok])
            jbs.append(JB[ok])
        if not ks:
            z = np.zeros(0, np.int64)
            return z, z, np.zeros(0), np.zeros((0, 2)), np.zeros(0)
        K = np.concatenate(ks)
        JB = np.concatenate(jbs)
        xa, xb, xp = x[self.ea[K]], x[self.eb[K]], x[self.bnd[JB]]
        ab = xb - xa
        L2 = np.maximum((ab * ab).sum(1), 1.0e-30)
        w = xp - xa
        tt = (w * ab).sum(1) / L2
        dvec = w - tt[:, None] * ab
        dvec = np.where((tt <= 0.0)[:, None], w, dvec)
        dvec = np.where((tt >= 1.0)[:, None], xp - xb, dvec)
        d = np.linalg.norm(dvec, axis=1)
        keep = d < self.reach
        return K[keep], JB[keep], tt[keep], dvec[keep], d[keep]

    # ---- one pass of the pair kernel (reference implicit/ipc_contact.py:175-296)
    def _run(self, s, mode, x, pdir, free, grad, diag, pHp, E, clr, alpha, mu, per_body, h2, floor):
        K, JB, tt, dvec, d = self.pairs(x)
        if K.size == 0:
            return
        a, b = self.ea[K], self.eb[K]
        e = self.e2env[K]
        pn = self.bnd[JB]
        dhat, kappa = self.dhat, self.kappa
        dlo = 1.0e-9 * dhat
        onset = 0.02 * dhat
        if mode in (1, 3):
            t_c = alpha[e].copy()
            np.minimum.at(clr, e, d)
            keep = np.maximum(self.gap * d, dhat - onset)
            inside = d < dhat
            keep = np.where(inside, d - np.minimum(2.0 * (dhat - d) + onset, 0.5 * d), keep)
            if mode == 3:
                keep = np.minimum(d, np.maximum(self.gap * d, floor))
            st = accd(x[pn], x[a], x[b], pdir[pn], pdir[a], pdir[b], keep, t_c)
            np.minimum.at(alpha, e, st)
        if mode == 3:
            return
        if per_body:
            mi, mj = self.mu_node[pn], self.mu_node[a]
            sm = mi + mj
            with np.errstate(divide="ignore", invalid="ignore"):
                mu_p = np.where(sm > 0.0, 2.0 * mi * mj / sm, 0.0)
        else:
            mu_p = np.full(K.size, mu)
        xa, xb, xp = x[a], x[b], x[pn]
        ab = xb - xa
        L2 = np.maximum((ab * ab).sum(1), 1.0e-30)
        L = np.sqrt(L2)
        w = xp - xa
        terms = []                     # (sel, coef, ip, ia, ib, wa, wb, d, n, ddd, e, mu)
        sel = (tt > 0.0) & (tt < 1.0) & (d < dhat)
        if sel.any():
            di = np.maximum(d[sel], dlo)
            n = dvec[sel] / di[:, None]
            ddd = np.zeros(sel.sum())
            if mode == 1:
                de = pdir[b[sel]] - pdir[a[sel]]
                dw = pdir[pn[sel]] - pdir[a[sel]]
                abs_, ws, Ls = ab[sel], w[sel], L[sel]
                c0 = cross2(abs_, ws)
                c1 = cross2(de, ws) + cross2(abs_, dw)
                c2 = 2.0 * cross2(de, dw)
                L1 = (abs_ * de).sum(1) / Ls
                Lq = ((de * de).sum(1) - L1 * L1) / Ls
                s0 = c0 / Ls
                s1 = (c1 - s0 * L1) / Ls
                ddd = (c2 - 2.0 * s1 * L1 - s0 * Lq) / Ls
                ddd = np.where(c0 < 0.0, -ddd, ddd)
            terms.append((np.ones(sel.sum()), pn[sel], a[sel], b[sel], 1.0 - tt[sel], tt[sel],
                          di, n, ddd, e[sel], mu_p[sel]))
        for end, cshare, onend, wvec, wa_, wb_ in ((a, self.ca[K], tt <= 0.0, w, 1.0, 0.0),
                                                   (b, self.cb[K], tt >= 1.0, xp - xb, 0.0, 1.0)):
            coef = -cshare + onend.astype(np.float64)
            dd = np.linalg.norm(wvec, axis=1)
            sel = (coef != 0.0) & (dd < dhat)
            if not sel.any():
                continue
            de_ = np.maximum(dd[sel], dlo)
            n = wvec[sel] / de_[:, None]
            ddd = np.zeros(sel.sum())
            if mode == 1:
                dw = pdir[pn[sel]] - pdir[end[sel]]
                dn = (dw * n).sum(1)
                ddd = ((dw * dw).sum(1) - dn * dn) / de_
            terms.append((coef[sel], pn[sel], a[sel], b[sel], np.full(sel.sum(), wa_),
                          np.full(sel.sum(), wb_), de_, n, ddd, e[sel], mu_p[sel]))
        for coef, ip, ia, ib, wa, wb, dd, n, ddd, ee, mup in terms:
            self._pair(mode, coef, ip, ia, ib, wa, wb, dd, n, ddd, ee, mup, x, s.xn, pdir,
                       s.epsv, s.inv_dt, h2, free, grad, diag, pHp, E)

    def _pair(self, mode, coef, ip, ia, ib, wa, wb, d, n, ddd, e, mu, x, xn, pdir, epsv, inv_dt,
              h2, free, grad, diag, pHp, E):
        """One barrier-and-friction term per row (reference implicit/ipc_contact.py:131-171)."""
        kappa, dhat = self.kappa, self.dhat
        bp = barrier_bp(d, dhat)
        f_n = -kappa * bp
        tn = np.stack([-n[:, 1], n[:, 0]], 1)
        u = (x[ip] - xn[ip]) - wa[:, None] * (x[ia] - xn[ia]) - wb[:, None] * (x[ib] - xn[ib])
        ut = (u * tn).sum(1)
        vt = np.abs(ut) * inv_dt
        kt = mu * f_n * inv_dt / np.maximum(vt, epsv)
        if mode == 0:
            bpp = barrier_bpp(d, dhat)
            g = h2 * (-f_n[:, None] * n + (mu * f_n * (ut * inv_dt / np.maximum(vt, epsv)))[:, None] * tn)
            hd = h2 * (kappa * bpp[:, None] * n * n + kt[:, None] * tn * tn)
            np.add.at(grad, ip, coef[:, None] * g * free[ip])
            np.add.at(diag, ip, coef[:, None] * hd)
            np.add.at(grad, ia, (-coef * wa)[:, None] * g * free[ia])
            np.add.at(diag, ia, (coef * wa * wa)[:, None] * hd)
            np.add.at(grad, ib, (-coef * wb)[:, None] * g * free[ib])
            np.add.at(diag, ib, (coef * wb * wb)[:, None] * hd)
        elif mode == 1:
            bpp = barrier_bpp(d, dhat)
            dq = pdir[ip] - wa[:, None] * pdir[ia] - wb[:, None] * pdir[ib]
            dn = (dq * n).sum(1)
            dt_ = (dq * tn).sum(1)
            val = np.maximum(kappa * (bpp * dn * dn + bp * ddd), 0.0)
            np.add.at(pHp, e, coef * h2 * (val + kt * dt_ * dt_))
        else:
            asl = epsv / inv_dt
            au = np.abs(ut)
            f0 = np.where(au < asl, au * au / (2.0 * asl), au - 0.5 * asl)
            np.add.at(E, e, coef * h2 * (kappa * barrier_b(d, dhat) + mu * f_n * f0))

    def _launch(self, s, mode, pdir, E=None, clr=None, alpha=None, grad=None, diag=None, free=None,
                friction=True, h2=None):
        self._run(s, mode, s.x, pdir, s.free if free is None else free,
                  s.grad if grad is None else grad, s.diagH if diag is None else diag, s.pHp,
                  s.Ener if E is None else E, s.clr_min if clr is None else clr,
                  s.alpha if alpha is None else alpha, self.mu if friction else 0.0,
                  self.per_body if friction else False, s.h2 if h2 is None else h2,
                  BC_FLOOR * self.dhat if mode == 3 else 0.0)

    # ---- term interface
    def grad_diag(self, s):
        self._launch(s, 0, s.p)

    def pHp(self, s):
        self.clr[:] = 1.0e30
        self._launch(s, 1, s.p, clr=self.clr)
        far_cap(s.p_inf, self.reach, s.alpha)

    def energy(self, s, buf):
        self._launch(s, 2, s.p, E=buf)

    def force(self, s, grad, diag, free, friction=True):
        self._launch(s, 0, s.p, grad=grad, diag=diag, free=free, friction=friction, h2=1.0)

    def ccd(self, s, pdir, p_inf, alpha):
        alpha[:] = 1.0
        self.clr[:] = 1.0e30
        self._launch(s, 3, pdir, clr=self.clr, alpha=alpha)
        far_cap(p_inf, self.reach, alpha)

    def min_distance(self, s):
        if self.q is None:
            self._alloc(s)
        self.q_inf[:] = 0.0
        self.ccd(s, self.q0, self.q_inf, self.beta)
        return self.clr.copy()

    # ---- the adjoint's tangent (no counterpart in the reference, whose adjoint refuses IPC)
    def terms(self, x):
        """The barrier terms the gradient pass evaluates at x, as arrays over terms: kind (0
        interior foot, 1 end a, 2 end b), coefficient, node, edge ends a and b. The same
        selection as _run: interior when 0 < t < 1 and d < dhat; an end when its
        vertex-correction coefficient is non-zero and the node is within dhat of it."""
        K, JB, tt, dvec, d = self.pairs(x)
        pn, a, b = self.bnd[JB], self.ea[K], self.eb[K]
        xp = x[pn]
        out = []
        sel = (tt > 0.0) & (tt < 1.0) & (d < self.dhat)
        out.append((np.zeros(sel.sum(), np.int64), np.ones(sel.sum()), pn[sel], a[sel], b[sel]))
        for kind, end, share, onend in ((1, a, self.ca[K], tt <= 0.0), (2, b, self.cb[K], tt >= 1.0)):
            coef = -share + onend.astype(np.float64)
            dd = np.linalg.norm(xp - x[end], axis=1)
            sel = (coef != 0.0) & (dd < self.dhat)
            out.append((np.full(sel.sum(), kind), coef[sel], pn[sel], a[sel], b[sel]))
        return tuple(np.concatenate([o[i] for o in out]) for i in range(5))

    def hvp(self, x, v, h2, y):
        """y += h2 (Hessian of the barrier energy) v at x, the terms and their coefficients frozen
        at x, friction left out, UNCLAMPED. Interior term: d = |s|, s = (e x w) / |e| with
        e = x_b - x_a, w = x_p - x_a, so grad s = (R(e) / L on p; -R(w) / L - s e / L^2 on b;
        minus their sum on a), R(u) = (-u_y, u_x), and the product is the directional derivative
        of grad s along v. End terms: d = |x_p - x_end|, Hess d = (I - n n^T) / d."""
        kind, coef, pn, a, b = self.terms(x)
        if kind.size == 0:
            return y
        dlo = 1.0e-9 * self.dhat
        R = lambda u: np.stack([-u[:, 1], u[:, 0]], 1)
        dot = lambda u, w: (u * w).sum(1)
        xp, xa, xb = x[pn], x[a], x[b]
        vp, va, vb = v[pn], v[a], v[b]
        yp, ya, yb = np.zeros_like(xp), np.zeros_like(xp), np.zeros_like(xp)
        it = kind == 0
        if it.any():
            e, w = (xb - xa)[it], (xp - xa)[it]
            de, dw = (vb - va)[it], (vp - va)[it]
            L2 = np.maximum(dot(e, e), 1.0e-30)
            L = np.sqrt(L2)
            c = cross2(e, w)
            d = np.maximum(np.abs(c) / L, dlo)
            sg = np.where(c < 0.0, -1.0, 1.0)
            gw = R(e) / L[:, None]
            ge = -R(w) / L[:, None] - (c / (L * L2))[:, None] * e
            sv = dot(gw, dw) + dot(ge, de)
            eL = dot(e, de)
            dc = cross2(de, w) + cross2(e, dw)
            dgw = R(de) / L[:, None] - (eL / (L * L2))[:, None] * R(e)
            dge = (-R(dw) / L[:, None] + (eL / (L * L2))[:, None] * R(w) - (dc / (L * L2))[:, None] * e
                   - (c / (L * L2))[:, None] * de + (3.0 * c * eL / (L2 * L2 * L))[:, None] * e)
            bp, bpp = barrier_bp(d, self.dhat), barrier_bpp(d, self.dhat)
            k = self.kappa * coef[it] * h2
            Hw = k[:, None] * (bpp[:, None] * sv[:, None] * gw + (bp * sg)[:, None] * dgw)
            He = k[:, None] * (bpp[:, None] * sv[:, None] * ge + (bp * sg)[:, None] * dge)
            yp[it] += Hw
            yb[it] += He
            ya[it] -= Hw + He
        for kk, xe, ve, ye in ((1, xa, va, ya), (2, xb, vb, yb)):
            m = kind == kk
            if not m.any():
                continue
            w = (xp - xe)[m]
            d = np.maximum(np.linalg.norm(w, axis=1), dlo)
            n = w / d[:, None]
            dq = (vp - ve)[m]
            dn = dot(dq, n)
            bp, bpp = barrier_bp(d, self.dhat), barrier_bpp(d, self.dhat)
            Hq = (self.kappa * coef[m] * h2)[:, None] * (bpp[:, None] * dn[:, None] * n +
                                                         (bp / d)[:, None] * (dq - dn[:, None] * n))
            yp[m] += Hq
            ye[m] -= Hq
        np.add.at(y, pn, yp)
        np.add.at(y, a, ya)
        np.add.at(y, b, yb)
        return y

    def barrier_grad_diag(self, s, x, h2, g, d):
        """The barrier part of the gradient pass (friction off) at x into g, d, nothing masked:
        the adjoint preconditioner's contact diagonal."""
        self._run(s, 0, x, np.zeros_like(x), np.ones_like(x), g, d, None, None, None, None, 0.0,
                  False, h2, 0.0)

    def _alloc(self, s):
        self.q = np.zeros((s.N, 2))
        self.q0 = np.zeros((s.N, 2))
        self.q_inf = np.zeros(s.B)
        self.beta = np.zeros(s.B)

    def _sweep(self, s):
        """Move the kinematic nodes toward their targets as far as the step limit allows
        (reference implicit/ipc_contact.py:308-331, 456-463)."""
        self.q_inf[:] = 0.0
        i = s.bc_idx
        dq = s.bc_target - s.x[i]
        self.q[i] = dq
        np.maximum.at(self.q_inf, s.node2env[i], np.linalg.norm(dq, axis=1))
        self.ccd(s, self.q, self.q_inf, self.beta)
        t = self.beta[s.node2env[i]]
        full = t >= 1.0
        xi = s.x[i] + t[:, None] * (s.bc_target - s.x[i])
        xi[full] = s.bc_target[full]
        s.x[i] = xi
        s.xhat[i] = xi

    def advance_bc(self, s):
        self.pending = False
        if s.n_bc == 0:
            return
        if self.q is None:
            self._alloc(s)
        self._sweep(s)
        self.pending = not bool((self.beta >= 1.0).all())

    def before_iteration(self, s, it):
        if not self.pending:
            return
        self._sweep(s)
        s.active[self.q_inf > 0.0] = 1
        if it % 8 == 7:
            self.pending = not bool((self.beta >= 1.0).all())


def far_cap(p_inf, reach, alpha):
    """alpha <= 0.45 reach / p_inf (reference implicit/ipc_contact.py:300-305)."""
    m = p_inf > 0.0
    alpha[m] = np.minimum(alpha[m], 0.45 * reach / p_inf[m])
