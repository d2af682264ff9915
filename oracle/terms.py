"""Element, barrier and collider math of the oracle (float64, vectorised numpy).

Follows reference core/math.py:35-144, implicit/kernels.py:73-530, core/sdf.py:45-306.
Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

BIG = 1.0e30
LOG_EPS = 1.0e-6


# ----------------------------------------------------------------------------- 2x2 algebra
def det2(F):
    return F[..., 0, 0] * F[..., 1, 1] - F[..., 0, 1] * F[..., 1, 0]


def inv2(F):
    """Inverse of 2x2 matrices by the adjugate (what a 2x2 inverse is)."""
    d = det2(F)
    out = np.empty_like(F)
    out[..., 0, 0] = F[..., 1, 1] / d
    out[..., 0, 1] = -F[..., 0, 1] / d
    out[..., 1, 0] = -F[..., 1, 0] / d
    out[..., 1, 1] = F[..., 0, 0] / d
    return out


def ddot(A, B):
    return (A * B).sum((-2, -1))


def ds_from_nodes(x, tri):
    """Edge vectors as COLUMNS: Ds = [x1-x0 | x2-x0] (reference core/math.py:35-39)."""
    p0, p1, p2 = x[tri[:, 0]], x[tri[:, 1]], x[tri[:, 2]]
    Ds = np.empty((tri.shape[0], 2, 2))
    Ds[:, :, 0] = p1 - p0
    Ds[:, :, 1] = p2 - p0
    return Ds


def dF_from_vec6(p, tri, B):
    """dF for a nodal direction p (reference core/math.py:68-73)."""
    return ds_from_nodes(p, tri) @ B


def dF_from_dir(n, c, B):
    """dF of a unit move of node n, component c (reference core/math.py:42-65)."""
    a, b = {0: (-1.0, -1.0), 1: (1.0, 0.0), 2: (0.0, 1.0)}[n]
    dDs = np.zeros(B.shape)
    dDs[:, c, 0] = a
    dDs[:, c, 1] = b
    return dDs @ B


def nh_logJ(F):
    return np.log(np.maximum(det2(F), LOG_EPS))


def nh_P(F, mu, lam):
    """First Piola stress mu F + (lam log J - mu) F^-T (reference core/math.py:76-82)."""
    G = np.swapaxes(inv2(F), -1, -2)
    c = lam * nh_logJ(F) - mu
    return mu[..., None, None] * F + c[..., None, None] * G


def nh_d2(F, dF, mu, lam):
    """d2Psi/dF2 [dF] = mu dF + lam (G:dF) G - (lam log J - mu) G dF^T G
    (reference core/math.py:85-94)."""
    G = np.swapaxes(inv2(F), -1, -2)
    c = lam * nh_logJ(F) - mu
    GdF = ddot(G, dF)
    GdFtG = G @ np.swapaxes(dF, -1, -2) @ G
    return (mu[..., None, None] * dF + (lam * GdF)[..., None, None] * G
            - c[..., None, None] * GdFtG)


def nh_psi(F, mu, lam):
    lj = nh_logJ(F)
    return 0.5 * mu * (ddot(F, F) - 2.0) - mu * lj + 0.5 * lam * lj * lj


# ----------------------------------------------------------------------------- elastic
def elastic_grad_diag(x, tri, B, area, mu, lam, h2, clamp, grad, diag):
    """grad += h2 A P B^T scattered to the 3 nodes; diag += h2 max(A dF:d2[dF], 0) for the
    6 element DOFs (reference implicit/kernels.py:86-128)."""
    F = ds_from_nodes(x, tri) @ B
    HH = area[:, None, None] * (nh_P(F, mu, lam) @ np.swapaxes(B, -1, -2))
    g1 = HH[:, :, 0]
    g2 = HH[:, :, 1]
    g0 = -(g1 + g2)
    contrib = np.stack([g0, g1, g2], 1) * h2                  # (T, 3, 2)
    np.add.at(grad, tri.reshape(-1), contrib.reshape(-1, 2))
    dvals = np.empty((tri.shape[0], 3, 2))
    for n in range(3):
        for c in range(2):
            dF = dF_from_dir(n, c, B)
            v = area * ddot(dF, nh_d2(F, dF, mu, lam))
            if clamp:
                v = np.maximum(v, 0.0)
            dvals[:, n, c] = h2 * v
    np.add.at(diag, tri.reshape(-1), dvals.reshape(-1, 2))


def elastic_pHp(x, p, tri, B, area, mu, lam, h2, clamp, tri2env, pHp):
    """pHp_env += h2 max(A dF:d2[dF], 0), dF from p (reference implicit/kernels.py:259-273)."""
    F = ds_from_nodes(x, tri) @ B
    dF = dF_from_vec6(p, tri, B)
    v = area * ddot(dF, nh_d2(F, dF, mu, lam))
    if clamp:
        v = np.maximum(v, 0.0)
    np.add.at(pHp, tri2env, h2 * v)


def elastic_energy(x, tri, B, area, mu, lam, h2, tri2env, E):
    F = ds_from_nodes(x, tri) @ B
    np.add.at(E, tri2env, h2 * area * nh_psi(F, mu, lam))


def elastic_hvp(x, v, tri, B, area, mu, lam, h2, y):
    """y += h2 (d2E/dx2) v, unclamped (reference implicit/adjoint.py:70-85)."""
    F = ds_from_nodes(x, tri) @ B
    dP = nh_d2(F, dF_from_vec6(v, tri, B), mu, lam)
    HH = area[:, None, None] * (dP @ np.swapaxes(B, -1, -2))
    h1, h2v = HH[:, :, 0], HH[:, :, 1]
    c = np.stack([-(h1 + h2v), h1, h2v], 1) * h2
    np.add.at(y, tri.reshape(-1), c.reshape(-1, 2))


def elastic_force(x, tri, B, area, mu, lam, h2, R):
    """R += h2 dE_el/dx (reference implicit/adjoint.py:54-66)."""
    F = ds_from_nodes(x, tri) @ B
    HH = area[:, None, None] * (nh_P(F, mu, lam) @ np.swapaxes(B, -1, -2))
    g1, g2 = HH[:, :, 0], HH[:, :, 1]
    c = np.stack([-(g1 + g2), g1, g2], 1) * h2
    np.add.at(R, tri.reshape(-1), c.reshape(-1, 2))


def inversion_alpha(x, p, tri, tri2env, shrink, alpha):
    """Smallest positive root of 2 area(s) = shrink * 2 area(0), atomic-min per env
    (reference implicit/kernels.py:304-341)."""
    e1 = x[tri[:, 1]] - x[tri[:, 0]]
    e2 = x[tri[:, 2]] - x[tri[:, 0]]
    q1 = p[tri[:, 1]] - p[tri[:, 0]]
    q2 = p[tri[:, 2]] - p[tri[:, 0]]
    a0 = cross2(e1, e2)
    a1 = cross2(e1, q2) + cross2(q1, e2)
    a2 = cross2(q1, q2)
    cc = a0 - shrink * a0
    s = np.full(tri.shape[0], BIG)
    lin = np.abs(a2) < 1.0e-24
    with np.errstate(divide="ignore", invalid="ignore"):
        sl = -cc / a1
        ok = lin & (np.abs(a1) > 1.0e-24) & (sl > 0.0)
        s[ok] = sl[ok]
        disc = a1 * a1 - 4.0 * a2 * cc
        quad = (~lin) & (disc >= 0.0)
        sq = np.sqrt(np.where(quad, disc, 0.0))
        r1 = (-a1 - sq) / (2.0 * a2)
        r2 = (-a1 + sq) / (2.0 * a2)
    r = np.full(tri.shape[0], BIG)
    m1 = quad & (r1 > 0.0)
    r[m1] = np.minimum(r[m1], r1[m1])
    m2 = quad & (r2 > 0.0)
    r[m2] = np.minimum(r[m2], r2[m2])
    s = np.where(quad, r, s)
    hit = s < BIG
    np.minimum.at(alpha, tri2env[hit], s[hit])


def cross2(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


# ----------------------------------------------------------------------------- barrier
def barrier_b(d, dhat):
    u = d - dhat
    return -(u * u) * np.log(d / dhat)


def barrier_bp(d, dhat):
    u = d - dhat
    return -2.0 * u * np.log(d / dhat) - (u * u) / d


def barrier_bpp(d, dhat):
    u = d - dhat
    return -2.0 * np.log(d / dhat) - 4.0 * u / d + (u * u) / (d * d)


# ----------------------------------------------------------------------------- colliders
def sdf_eval(kind, cx, cy, rx, ry, rref, X):
    """(phi, grad phi (n,2), hess phi (n,2,2)) of a collider level set at points X
    (reference core/sdf.py:45-137)."""
    n = X.shape[0]
    g = np.zeros((n, 2))
    H = np.zeros((n, 2, 2))
    if kind == 0:
        diff = X - np.array([cx, cy])
        dc = np.linalg.norm(diff, axis=1)
        nh = diff / dc[:, None]
        d = dc - rx
        g = nh
        H = (np.eye(2)[None] - nh[:, :, None] * nh[:, None, :]) * (1.0 / dc)[:, None, None]
    elif kind == 1:
        ux = (X[:, 0] - cx) / rx
        uy = (X[:, 1] - cy) / ry
        u = ux * ux + uy * uy
        su = np.sqrt(np.maximum(u, 1.0e-18))
        w = np.stack([(X[:, 0] - cx) / (rx * rx), (X[:, 1] - cy) / (ry * ry)], 1)
        d = rref * (su - 1.0)
        g = (rref / su)[:, None] * w
        D = np.diag([1.0 / (rx * rx), 1.0 / (ry * ry)])
        H = (rref / su)[:, None, None] * (D[None] - (1.0 / u)[:, None, None] * w[:, :, None] * w[:, None, :])
    elif kind == 4:
        nvec = np.array([rx, ry])
        d = (X - np.array([cx, cy])) @ nvec
        g = np.broadcast_to(nvec, (n, 2)).copy()
    elif kind == 2:
        r = min(max(rref, 0.0), min(rx, ry))
        sx = np.where(X[:, 0] < cx, -1.0, 1.0)
        sy = np.where(X[:, 1] < cy, -1.0, 1.0)
        qx = sx * (X[:, 0] - cx) - (rx - r)
        qy = sy * (X[:, 1] - cy) - (ry - r)
        mx, my = np.maximum(qx, 0.0), np.maximum(qy, 0.0)
        ml = np.sqrt(mx * mx + my * my)
        d = ml + np.minimum(np.maximum(qx, qy), 0.0) - r
        out = ml > 1.0e-12
        with np.errstate(divide="ignore", invalid="ignore"):
            gx, gy = mx / ml, my / ml
        corner = out & (qx > 0.0) & (qy > 0.0)
        g[out, 0] = (sx * gx)[out]
        g[out, 1] = (sy * gy)[out]
        off = -gx * gy / ml
        H[corner, 0, 0] = (gy * gy / ml)[corner]
        H[corner, 0, 1] = (sx * sy * off)[corner]
        H[corner, 1, 0] = (sx * sy * off)[corner]
        H[corner, 1, 1] = (gx * gx / ml)[corner]
        inside = ~out
        px = inside & (qx > qy)
        py = inside & ~(qx > qy)
        g[px, 0] = sx[px]
        g[py, 1] = sy[py]
    elif kind == 3:
        r3 = min(max(rref, 0.0), rx * 0.5)
        rin = rx - r3
        c0 = np.array([cx, cy])
        a0 = 1.5707963267948966
        V = [c0 + rin * np.array([np.cos(a0 + k * 2.0943951023931953), np.sin(a0 + k * 2.0943951023931953)])
             for k in range(3)]
        best = None
        for (va, vb) in ((V[0], V[1]), (V[1], V[2]), (V[2], V[0])):
            ab = vb - va
            t = ((X - va) @ ab) / max(ab @ ab, 1.0e-30)
            tc = np.clip(t, 0.0, 1.0)
            q = va + tc[:, None] * ab
            dd = np.linalg.norm(X - q, axis=1)
            isv = (t <= 0.0) | (t >= 1.0)
            if best is None:
                best = [dd, q, isv]
            else:
                upd = dd < best[0]
                best[0] = np.where(upd, dd, best[0])
                best[1] = np.where(upd[:, None], q, best[1])
                best[2] = np.where(upd, isv, best[2])
        dm, qm, sm = best
        e = [V[1] - V[0], V[2] - V[1], V[0] - V[2]]
        k = [e[i][0] * (X[:, 1] - V[i][1]) - e[i][1] * (X[:, 0] - V[i][0]) for i in range(3)]
        sgn = np.where((k[0] >= 0) & (k[1] >= 0) & (k[2] >= 0), -1.0, 1.0)
        dl = np.maximum(dm, 1.0e-30)
        d = sgn * dm - r3
        g = (sgn / dl)[:, None] * (X - qm)
        rc = sm & (sgn > 0)
        gg = g[rc]
        H[rc] = (np.eye(2)[None] - gg[:, :, None] * gg[:, None, :]) * (1.0 / dl[rc])[:, None, None]
    else:
        raise ValueError(f"unknown collider kind {kind}")
    return d, g, H


def _gate(d, dhat):
    return (d > 0.0) & (d < dhat)


def collider_grad_diag(col, x, bnd, dhat, kappa, h2, clamp, grad, diag):
    """Barrier gradient and diagonal against one collider. Disc: implicit/kernels.py:132-157;
    other shapes: core/sdf.py:141-154."""
    X = x[bnd]
    if col["kind"] == 0:
        diff = X - np.array([col["cx"], col["cy"]])
        dc = np.linalg.norm(diff, axis=1)
        d = dc - col["r"]
        m = _gate(d, dhat)
        nh = diff[m] / dc[m, None]
        bp, bpp = barrier_bp(d[m], dhat), barrier_bpp(d[m], dhat)
        np.add.at(grad, bnd[m], (h2 * kappa * bp)[:, None] * nh)
        tang = bp / dc[m]
        dg = kappa * (bpp[:, None] * nh * nh + tang[:, None] * (1.0 - nh * nh))
    else:
        d, g, H = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
        m = _gate(d, dhat)
        bp, bpp = barrier_bp(d[m], dhat), barrier_bpp(d[m], dhat)
        np.add.at(grad, bnd[m], (h2 * kappa * bp)[:, None] * g[m])
        Hd = np.stack([H[m, 0, 0], H[m, 1, 1]], 1)
        dg = kappa * (bpp[:, None] * g[m] * g[m] + bp[:, None] * Hd)
    if clamp:
        dg = np.maximum(dg, 0.0)
    np.add.at(diag, bnd[m], h2 * dg)


def collider_pHp(col, x, p, bnd, b2e, dhat, kappa, h2, clamp, pHp, clr_min):
    """Barrier p^T H p and the clearance minimum (implicit/kernels.py:277-299, core/sdf.py:158-171)."""
    X = x[bnd]
    if col["kind"] == 0:
        diff = X - np.array([col["cx"], col["cy"]])
        dc = np.linalg.norm(diff, axis=1)
        d = dc - col["r"]
        np.minimum.at(clr_min, b2e, d)
        m = _gate(d, dhat)
        nh = diff[m] / dc[m, None]
        pp = p[bnd[m]]
        pn = (pp * nh).sum(1)
        pt = pp - pn[:, None] * nh
        val = kappa * (barrier_bpp(d[m], dhat) * pn * pn + (barrier_bp(d[m], dhat) / dc[m]) * (pt * pt).sum(1))
    else:
        d, g, H = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
        np.minimum.at(clr_min, b2e, d)
        m = _gate(d, dhat)
        pp = p[bnd[m]]
        gp = (g[m] * pp).sum(1)
        pHp_ = (pp * np.einsum("nij,nj->ni", H[m], pp)).sum(1)
        val = kappa * (barrier_bpp(d[m], dhat) * gp * gp + barrier_bp(d[m], dhat) * pHp_)
    if clamp:
        val = np.maximum(val, 0.0)
    np.add.at(pHp, b2e[m], h2 * val)


def collider_energy(col, x, bnd, b2e, dhat, kappa, h2, E):
    X = x[bnd]
    if col["kind"] == 0:
        d = np.linalg.norm(X - np.array([col["cx"], col["cy"]]), axis=1) - col["r"]
    else:
        d = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)[0]
    m = _gate(d, dhat)
    np.add.at(E, b2e[m], h2 * kappa * barrier_b(d[m], dhat))


def _friction_frame(col, x, xn, bnd, dhat, kappa, inv_dt):
    """(gate, f_n, nh, vt, |vt|) of the lagged friction at boundary nodes (disc:
    implicit/kernels.py:661-690; level set: core/sdf.py:201-210)."""
    X = x[bnd]
    if col["kind"] == 0:
        diff = X - np.array([col["cx"], col["cy"]])
        dc = np.linalg.norm(diff, axis=1)
        d = dc - col["r"]
        m = _gate(d, dhat)
        nh = diff[m] / dc[m, None]
        f_n = kappa * np.abs(barrier_bp(d[m], dhat))
    else:
        d, g, _ = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
        m = _gate(d, dhat)
        gn = np.linalg.norm(g[m], axis=1)
        nh = g[m] / np.maximum(gn, 1.0e-30)[:, None]
        f_n = kappa * np.abs(barrier_bp(d[m], dhat)) * gn
    v = (x[bnd[m]] - xn[bnd[m]]) * inv_dt
    vt = v - (v * nh).sum(1)[:, None] * nh
    return m, f_n, nh, vt, np.linalg.norm(vt, axis=1)


def _free_gate(free, bnd):
    fr = free[bnd]
    return ~((fr[:, 0] <= 0.5) & (fr[:, 1] <= 0.5))


def friction_grad_diag(col, x, xn, bnd, dhat, kappa, mu, epsv, inv_dt, h2, free, grad, diag):
    fg = _free_gate(free, bnd)
    b = bnd[fg]
    m, f_n, nh, vt, vtm = _friction_frame(col, x, xn, b, dhat, kappa, inv_dt)
    n = b[m]
    fr = free[n]
    tdir = vt / np.maximum(vtm, epsv)[:, None]
    np.add.at(grad, n, (h2 * mu * f_n)[:, None] * tdir * fr)
    kt = h2 * mu * f_n * inv_dt / np.maximum(vtm, epsv)
    th = np.stack([-nh[:, 1], nh[:, 0]], 1)
    np.add.at(diag, n, kt[:, None] * th * th * fr)


def friction_pHp(col, x, xn, p, bnd, b2e, dhat, kappa, mu, epsv, inv_dt, h2, free, pHp):
    fg = _free_gate(free, bnd)
    b, be = bnd[fg], b2e[fg]
    m, f_n, nh, vt, vtm = _friction_frame(col, x, xn, b, dhat, kappa, inv_dt)
    n = b[m]
    kt = h2 * mu * f_n * inv_dt / np.maximum(vtm, epsv)
    th = np.stack([-nh[:, 1], nh[:, 0]], 1)
    pt = (p[n] * free[n] * th).sum(1)
    np.add.at(pHp, be[m], kt * pt * pt)


def friction_energy(col, x, xn, bnd, b2e, dhat, kappa, mu, epsv, inv_dt, h2, free, E):
    fg = _free_gate(free, bnd)
    b, be = bnd[fg], b2e[fg]
    m, f_n, nh, vt, vtm = _friction_frame(col, x, xn, b, dhat, kappa, inv_dt)
    ut = vtm / inv_dt
    a = epsv / inv_dt
    f0 = np.where(ut < a, ut * ut / (2.0 * a), ut - 0.5 * a)
    np.add.at(E, be[m], h2 * mu * f_n * f0)


def collider_ccd(col, x, p, bnd, b2e, eta, alpha):
    """Step limit against a collider: exact quadratic time of impact for a disc
    (implicit/kernels.py:345-371), conservative linear bound for a level set (core/sdf.py:298-306)."""
    X = x[bnd]
    P = p[bnd]
    if col["kind"] == 0:
        o = X - np.array([col["cx"], col["cy"]])
        a = (P * P).sum(1)
        bq = 2.0 * (o * P).sum(1)
        cq = (o * o).sum(1) - col["r"] * col["r"]
        with np.errstate(divide="ignore", invalid="ignore"):
            disc = bq * bq - 4.0 * a * cq
            ok = (a >= 1.0e-24) & (disc >= 0.0)
            sq = np.sqrt(np.where(ok, disc, 0.0))
            r1 = (-bq - sq) / (2.0 * a)
            r2 = (-bq + sq) / (2.0 * a)
        s = np.full(len(bnd), BIG)
        s = np.where(ok & (r1 > 0.0), np.minimum(s, r1), s)
        s = np.where(ok & (r2 > 0.0), np.minimum(s, r2), s)
        hit = s < BIG
        np.minimum.at(alpha, b2e[hit], eta * s[hit])
    else:
        d, g, _ = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
        gp = (g * P).sum(1)
        m = ~((gp >= 0.0) | (d <= 0.0))
        np.minimum.at(alpha, b2e[m], eta * d[m] / (-gp[m]))


def collider_hvp(col, x, v, bnd, dhat, kappa, h2, y):
    """y += h2 H_barrier v at a frozen gate (implicit/adjoint.py:89-112)."""
    X = x[bnd]
    d, g, H = sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
    m = _gate(d, dhat)
    vv = v[bnd[m]]
    gv = (g[m] * vv).sum(1)
    out = h2 * kappa * (barrier_bpp(d[m], dhat)[:, None] * gv[:, None] * g[m]
                        + barrier_bp(d[m], dhat)[:, None] * np.einsum("nij,nj->ni", H[m], vv))
    np.add.at(y, bnd[m], out)
