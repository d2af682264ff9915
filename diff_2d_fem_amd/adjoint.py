"""Adjoint of the implicit step: dJ/d(density) through the implicit function theorem
(reference implicit/adjoint.py:1-399, AdjointPNCG2D).

Density enters as a per-element multiplier of the full-density moduli,
mu_e = E(rho_e) mu0_e, lam_e = E(rho_e) lam0_e with E(rho) = Emin + rho (1 - Emin). A
trajectory x_0 .. x_T is differentiated by the damped backward-Euler recursion

    H_k lam_k = dJ/dx_k + (1 + drag) M lam_{k+1} - drag M lam_{k+2}
    dJ/drho_e = sum_k -(1 - Emin) lam_k . [h^2 f_unit_e(x_k)]

with H_k = M + h^2 (K_elastic + K_contact + K_pressure) at x_k, restricted to the free degrees of
freedom, the contact active set frozen at x_k. The operator holds inertia, elasticity, the
colliders' barriers, with use_self the "node" contact's penalty between bodies, and — beyond
the reference, whose adjoint refuses them:
* pressure chambers, whose tangent is exact (the pressure gradient is linear in x); the
  recursion then also gives dJ/dp per chamber and step (dJdp_steps):
  dJ/dp_c,k = -lam_k . dR_k/dp_c;
* the IPC barrier between bodies (use_ipc, on by default when the model has IPC contact): the
  Hessian of every node-edge barrier term the gradient pass evaluates at x_k, the terms and
  their vertex-correction coefficients frozen there;
* affine bodies: their slaved nodes follow x_i = J_i q, so the step's optimality conditions are
  those of the energy in (free soft nodes, q). Each solve is P^T H P lam_r = P^T rhs with P that
  map, and lam = P lam_r feeds dJ/drho and the recursion. A region's own elements carry their
  moduli in H, so at the slaved positions their elastic Hessian is that of the region energy
  in q, and dJ/drho covers them.
Friction (colliders: no tangent, as in the reference; between bodies: refused), tendons, springs
and active strain have no tangent in it (transient_adjoint refuses a model with any of them
except collider friction, as the reference does). Rigid modes only precondition the forward
solve and do not enter.

Everything runs on the GPU: the Hessian-vector products, dJ/drho and dJ/dp are HIP kernels
(d2f_elastic_hvp_f64, d2f_collider_hvp_f64, d2f_node_hvp_f64, d2f_ipc_hvp_f64,
d2f_pressure_hvp_f64, d2f_elastic_drho_f64, d2f_pressure_dp_f64), the affine reduction is
d2f_affine_expand_f64 / d2f_affine_reduce_f64, and each adjoint solve is a preconditioned MINRES
with per-scene Krylov scalars on the device (d2f_minres_step_f64): the scenes of a batch do not
couple through shared scalars, and the host reads only the scenes' status flags, every
`check_every` iterations.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib

F64 = torch.float64


class AdjointPNCG2D:
    """dJ/drho of a DeviceSolver's implicit steps (reference implicit/adjoint.py:181-399)."""

    def __init__(self, solver, Emin=1e-3):
        self.s = solver
        self.dev = solver.dev
        me = solver.me
        self.Emin = float(Emin)
        self.mu0 = torch.as_tensor(np.asarray(me.mu, np.float64), device=self.dev)
        self.lam0 = torch.as_tensor(np.asarray(me.lam, np.float64), device=self.dev)
        self.N, self.n_tri = solver.N, solver.n_tri
        self.h2 = float(solver.h2)
        self.free_np = np.asarray(solver.free_np, np.float64)          # (N, 2): 1 free, 0 held
        self._free = torch.as_tensor(self.free_np, device=self.dev)
        self.refresh_mass()
        self._y = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        self._drho = torch.zeros(self.n_tri, dtype=F64, device=self.dev)
        cols = solver.colliders
        self._cols = (_lib.Collider * max(1, len(cols)))()
        for i, c in enumerate(cols):
            self._cols[i].kind = int(c["kind"])
            for k in ("cx", "cy", "rx", "ry", "rref", "r", "mu"):
                setattr(self._cols[i], k, float(c[k]))
        self._n_col = len(cols)
        self._tiles = solver.elem_tiles_struct()
        # moduli of the operator: the solver's, with each rigid region's own elements (zero in
        # the solver's nodal passes) carrying the moduli its region energy sums
        mu_el, lam_el = np.asarray(solver.mu_np, np.float64).copy(), np.asarray(solver.lam_np, np.float64).copy()
        self.affine = solver.affine
        if self.affine is not None:
            for el in self.affine["group_elems"]:
                mu_el[el] = np.asarray(me.mu, np.float64)[el]
                lam_el[el] = np.asarray(me.lam, np.float64)[el]
            self._setup_reduction()
        self._mu_el, self._lam_el = self._t(mu_el), self._t(lam_el)
        self._pairs = None                   # IPC pair list at the state being solved
        self._n_pairs = 0

    def _setup_reduction(self):
        """Index maps of the affine reduction: slave_of[i] (entry of node i in the an_* lists, or
        -1), three slot nodes per group, the reduced free mask and the soft-node mass."""
        a, s = self.affine, self.s
        node = a["node_np"]
        slave_of = np.full(self.N, -1, np.int32)
        slave_of[node] = np.arange(len(node), dtype=np.int32)
        first = a["first"]
        slot = np.stack([node[first[:-1] + k] for k in range(3)], 1).reshape(-1).astype(np.int32)
        fm = self.free_np.copy()
        fm[node] = 0.0
        fm[slot] = 1.0
        I32 = torch.int32
        self._slave_of = torch.as_tensor(slave_of, device=self.dev)
        self._slot = torch.as_tensor(slot, device=self.dev)
        self._an_first = torch.as_tensor(np.asarray(first, np.int32), device=self.dev).to(I32)
        self._fmask_red = self._t(fm)
        self._slaved = torch.as_tensor(a["slaved"], device=self.dev)
        self._n_groups = int(a["n"])
        self._vfull = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        self._kfull = torch.zeros((self.N, 2), dtype=F64, device=self.dev)

    # ---- plumbing
    def _t(self, a):
        if torch.is_tensor(a):
            return a.to(device=self.dev, dtype=F64).contiguous()
        return torch.as_tensor(np.asarray(a, np.float64), device=self.dev).contiguous()

    @staticmethod
    def _p(t):
        return C.c_void_p(t.data_ptr())

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def refresh_mass(self):
        """Re-read the solver's nodal mass (call after changing solver.mass)."""
        self.mass_np = np.asarray(self.s.mass.cpu().numpy(), np.float64)
        self._mass2 = self.s.mass.to(F64)[:, None]
        self._mass_full = self.s.mass.to(F64).clone()
        return self.mass_np

    # ---- density (reference implicit/adjoint.py:41-50, 225-228)
    def set_density(self, rho_e):
        rho = np.asarray(rho_e, np.float64).reshape(-1)
        if rho.shape[0] != self.n_tri:
            raise ValueError(f"density has {rho.shape[0]} entries, the mesh {self.n_tri} elements")
        E = self.Emin + rho * (1.0 - self.Emin)
        mu_el = E * np.asarray(self.s.me.mu, np.float64)
        lam_el = E * np.asarray(self.s.me.lam, np.float64)
        self._mu_el, self._lam_el = self._t(mu_el), self._t(lam_el)
        if self.affine is not None:          # regions: summed moduli, own elements zero in the solver
            self.s.set_affine_moduli(mu_el, lam_el)
        else:
            self.s.mu_np, self.s.lam_np = mu_el, lam_el
            self.s.refresh_moduli()

    # ---- operator H v = M v + h^2 (K_elastic + K_contact) v (reference :231-256)
    def _node_on(self, use_self):
        return bool(use_self) and self.s.node is not None

    def _ipc_on(self, use_ipc):
        return self.s.ipc is not None and (use_ipc is None or bool(use_ipc))

    def _set_pairs(self, xw):
        """The IPC pair list at xw (d2f_ipc_pairs), kept for the solve's products."""
        s = self.s
        L = _lib.lib()
        cap = max(int(getattr(self, "_pair_cap", 0)), 16 * max(s.n_e, 1))
        while True:
            if self._pairs is None or self._pairs.numel() < 2 + 2 * cap:
                self._pairs = torch.zeros(2 + 2 * cap, dtype=torch.int32, device=self.dev)
            sc, q = s.scene_struct("plain"), s.params_struct()
            sc.x = self._p(xw)
            _lib.check(L.d2f_ipc_pairs(C.byref(sc), C.byref(q), self._p(self._pairs), cap, self._stream()),
                       "d2f_ipc_pairs")
            n = int(self._pairs[0].item())
            if n <= cap:
                break
            cap = 2 * n
        self._pair_cap, self._n_pairs = cap, n

    def _kv_calls(self, xw, vw, y, use_disc=False, use_sdf=False, use_self=False, use_ipc=None):
        """The entry-point calls that form y += h^2 K v at xw (elastic, selected colliders, node
        contact, IPC barrier between bodies with the pairs of the last _set_pairs, pressure), as
        (function, bound ctypes arguments, name): the Krylov loop binds them once per solve."""
        L = _lib.lib()
        s = self.s
        P, st = self._p, self._stream()
        X, V, Y = P(xw), P(vw), P(y)
        pt = self._pressure_tiles() if s.pressure is not None else None
        if pt is not None:                   # elasticity and pressure in one tiled pass
            calls = [(L.d2f_elastic_pressure_hvp_f64,
                      (X, V, P(s.Bm), P(s.area), P(self._mu_el), P(self._lam_el), self.n_tri, self.h2,
                       C.byref(self._tiles), P(pt["pe_off"]), P(pt["pe_loc"]), P(pt["pe_ch"]), P(s.ch_p), Y, st),
                      "d2f_elastic_pressure_hvp_f64")]
        else:
            calls = [(L.d2f_elastic_hvp_f64, (X, V, P(s.tri), P(s.Bm), P(s.area), P(self._mu_el), P(self._lam_el),
                                              self.n_tri, self.h2, C.byref(self._tiles), Y, st),
                      "d2f_elastic_hvp_f64")]
        which = 2 if (use_disc and use_sdf) else (0 if use_disc else 1)
        if (use_disc or use_sdf) and self._n_col:
            calls.append((L.d2f_collider_hvp_f64, (X, V, P(s.bnd), s.n_b, self._cols, self._n_col, which,
                                                   float(s.cfg.ipc_d_hat), float(s.cfg.obstacle_stiffness), self.h2,
                                                   Y, st), "d2f_collider_hvp_f64"))
        if self._node_on(use_self):           # reference :239-245, the pairs frozen at x
            nd = s.node
            calls.append((L.d2f_node_hvp_f64, (X, V, P(s.bnd), P(s.bnd_start), s.B, P(s.excl_bits),
                                               P(s.excl_bits_off), nd["r"], nd["k"], self.h2, Y, st),
                          "d2f_node_hvp_f64"))
        if self._ipc_on(use_ipc) and self._n_pairs:
            ip = s.ipc
            calls.append((L.d2f_ipc_hvp_f64, (X, V, P(self._pairs), self._n_pairs, P(s.ea), P(s.eb), P(s.ca),
                                              P(s.cb), ip["dhat"], ip["kappa"], self.h2, Y, st), "d2f_ipc_hvp_f64"))
        if s.pressure is not None and pt is None:
            calls.append((L.d2f_pressure_hvp_f64, (V, P(s.pe_a), P(s.pe_b), P(s.pe_ch), P(s.ch_p),
                                                   s.pressure["n_pe"], self.h2, Y, st), "d2f_pressure_hvp_f64"))
        return calls

    def _pressure_tiles(self):
        """The chamber edges dealt to the element tiles (layout.pressure_tiles) on the device,
        built once; None when the tiles are not in use or an edge is no triangle's side."""
        if getattr(self, "_ptiles", None) is None:
            from .layout import ELEM_NODES_MAX, pressure_tiles
            s, et = self.s, self.s.el_tiles
            pt = None
            if s.n_tri and et["max_nodes"] <= ELEM_NODES_MAX:
                pt = pressure_tiles(s.tri.cpu().numpy(), et["tri_local"].cpu().numpy(), s.pe_a.cpu().numpy(),
                                    s.pe_b.cpu().numpy(), s.pe_ch.cpu().numpy(), self.N)
            self._ptiles = False if pt is None else {
                k: torch.as_tensor(v.view(np.int32) if v.dtype == np.uint32 else v, device=self.dev)
                for k, v in pt.items()}
        return self._ptiles or None

    def _kv(self, xw, vw, y, use_disc=False, use_sdf=False, use_self=False, use_ipc=None):
        """y += h^2 K v at xw (the calls of _kv_calls)."""
        for fn, args, name in self._kv_calls(xw, vw, y, use_disc, use_sdf, use_self, use_ipc):
            _lib.check(fn(*args), name)
        return y

    # ---- the affine reduction (reduced vectors: soft entries in place, q in three slot nodes)
    def _expand(self, vr, out):
        L = _lib.lib()
        s = self.s
        _lib.check(L.d2f_affine_expand_f64(self._p(vr), self._p(self._slave_of), self._p(s.an_grp),
                                           self._p(s.an_w), self._p(self._slot), self.N, self._p(out),
                                           self._stream()), "d2f_affine_expand_f64")
        return out

    def _reduce(self, yf, out, vf=None, square=False):
        L = _lib.lib()
        s = self.s
        _lib.check(L.d2f_affine_reduce_f64(self._p(yf), None if vf is None else self._p(vf),
                                           self._p(self._mass_full), self._p(self._slave_of),
                                           self._p(s.an_node), self._p(s.an_w), self._p(self._an_first),
                                           self._p(self._slot), self.N, self._n_groups, int(bool(square)),
                                           self._p(out), self._stream()), "d2f_affine_reduce_f64")
        return out

    def _kv_reduced(self, xw, vr, out, **use):
        """out = P^T (h^2 K + M_slaved) P vr: the reduced operator's part beyond the soft nodes'
        diagonal mass (the Krylov step adds mass v with the slaved nodes' mass zero)."""
        vf, kf = self._vfull, self._kfull
        self._expand(vr, vf)
        kf.zero_()
        self._kv(xw, vf, kf, **use)
        return self._reduce(kf, out, vf=vf)

    def _hvp(self, xw, vw, use_disc=False, use_sdf=False, include_mass=True, use_self=False, use_ipc=None):
        y = self._y
        y.zero_()
        self._kv(xw, vw, y, use_disc, use_sdf, use_self, use_ipc)
        return self._mass2 * vw + y if include_mass else y.clone()

    def apply_H(self, xstar, v, use_self=False, use_disc=False, use_sdf=False, include_mass=True, use_ipc=None):
        """H v at x* as an (N, 2) array (the reference's matrix-free product; node space, the rigid
        regions' elements with their moduli)."""
        xw = self._t(xstar)
        if self._ipc_on(use_ipc):
            self._set_pairs(xw)
        return self._hvp(xw, self._t(v), use_disc, use_sdf, include_mass, use_self=use_self,
                         use_ipc=use_ipc).cpu().numpy()

    def spd_probe(self, xstar, use_self=False, include_mass=True, n=16, seed=0):
        """min and max Rayleigh quotient p^T H p / p^T p over random free directions."""
        rng = np.random.default_rng(seed)
        xw = self._t(xstar)
        vals = []
        for _ in range(n):
            p = self._t(rng.standard_normal((self.N, 2))) * self._free
            Hp = self._hvp(xw, p, include_mass=include_mass, use_self=use_self) * self._free
            vals.append(float((p * Hp).sum() / (p * p).sum()))
        return min(vals), max(vals)

    # ---- Jacobi preconditioner diag(M + h^2 (K_elastic + contact)) (reference :356-372)
    def _diag_precond(self, xw, use_disc=False, use_sdf=False, use_self=False, use_ipc=None):
        L = _lib.lib()
        s = self.s
        g = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        d = self._mass2.expand(self.N, 2).clone()
        _lib.check(L.d2f_elastic_grad_diag_f64(self._p(xw), self._p(s.tri), self._p(s.Bm), self._p(s.area),
                                               self._p(self._mu_el), self._p(self._lam_el), self.n_tri,
                                               self.h2, 1, C.byref(self._tiles), self._p(g), self._p(d),
                                               self._stream()), "elastic grad/diag")
        if self._ipc_on(use_ipc) and self._n_pairs:
            ip = s.ipc
            _lib.check(L.d2f_ipc_grad_diag_f64(self._p(xw), self._p(self._pairs), self._n_pairs, self._p(s.ea),
                                               self._p(s.eb), self._p(s.ca), self._p(s.cb), ip["dhat"],
                                               ip["kappa"], self.h2, self._p(g), self._p(d), self._stream()),
                       "d2f_ipc_grad_diag_f64")
        if (use_disc or use_sdf) and self._n_col:
            which = 2 if (use_disc and use_sdf) else (0 if use_disc else 1)
            _lib.check(L.d2f_collider_grad_diag_f64(self._p(xw), self._p(s.bnd), s.n_b, self._cols,
                                                    self._n_col, which, float(s.cfg.ipc_d_hat),
                                                    float(s.cfg.obstacle_stiffness), self.h2, 1,
                                                    self._p(g), self._p(d), self._stream()),
                       "collider grad/diag")
        if self._node_on(use_self):           # reference :332-335: clamped, as in the solve
            nd = s.node
            _lib.check(L.d2f_node_grad_diag_f64(self._p(xw), self._p(s.bnd), self._p(s.bnd_start), s.B,
                                                self._p(s.excl_bits), self._p(s.excl_bits_off), nd["r"], nd["k"],
                                                self.h2, 1, self._p(g), self._p(d), self._stream()),
                       "d2f_node_grad_diag_f64")
        return torch.clamp(d, min=1e-30)

    # ---- Krylov solves on the free degrees of freedom
    def _workspace(self):
        ws = getattr(self, "_ws", None)
        if ws is None:
            z = lambda: torch.zeros((self.N, 2), dtype=F64, device=self.dev)
            ws = {k: z() for k in ("dinv", "b", "x", "v", "r1", "r2", "w1", "w2", "Kv")}
            ws["scal"] = torch.zeros(8 * self.s.B, dtype=F64, device=self.dev)
            ws["status"] = torch.zeros(2 * self.s.B, dtype=torch.int32, device=self.dev)
            ws["mass"] = self.s.mass.to(F64).clone()            # owned: the solve masks it
            self._ws = ws
        return ws

    def _minres(self, xw, b, dinv, rtol, maxiter, check_every=16, **use):
        """Preconditioned MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12, 1975) for every
        scene at once, each with its own Krylov scalars, on the device (d2f_minres_*). The
        operator H v = free (M v + h^2 K v) is symmetric, possibly indefinite; the preconditioner
        is the diagonal dinv (zero on held degrees of freedom, so iterates stay on the free ones).
        A scene stops when its preconditioned residual norm falls to rtol times that of its b.
        Returns (x, per-scene status (0 running at the limit, 1 converged, 2 iteration limit,
        -1 breakdown), per-scene iterations)."""
        L = _lib.lib()
        ws = self._workspace()
        red = self.affine is not None
        ws["mass"].copy_(self.s.mass)
        if red:                              # the slaved nodes' inertia is in the reduced product
            ws["mass"].masked_fill_(self._slaved, 0.0)
        ws["b"].copy_(b)
        ws["dinv"].copy_(dinv)
        m = _lib.Minres()
        m.B, m.maxiter, m.rtol = self.s.B, int(maxiter), float(rtol)
        fmask = self._fmask_red if red else self._free
        m.node_start, m.mass, m.fmask = self._p(self.s.node_start), self._p(ws["mass"]), self._p(fmask)
        for k in ("dinv", "b", "x", "v", "r1", "r2", "w1", "w2", "Kv", "scal", "status"):
            setattr(m, k, self._p(ws[k]))
        st = self._stream()
        _lib.check(L.d2f_minres_begin_f64(C.byref(m), st), "d2f_minres_begin_f64")
        step, mref = L.d2f_minres_step_f64, C.byref(m)

        def chunk(stream, calls):
            # check_every iterations: the products add into Kv (left cleared by the previous begin
            # or step), then the Krylov step; a scene past its limit or converged is left alone
            for _ in range(check_every):
                if calls is None:
                    self._kv_reduced(xw, ws["v"], ws["Kv"], **use)
                else:
                    for fn, args, name in calls:
                        rc = fn(*args)
                        if rc:
                            _lib.check(rc, name)
                rc = step(mref, stream)
                if rc:
                    _lib.check(rc, "d2f_minres_step_f64")
        calls = None if red else self._kv_calls(xw, ws["v"], ws["Kv"], **use)
        done = 0
        while done < maxiter:
            chunk(st, calls)
            done += check_every
            if not bool((ws["status"][0::2] == 0).any()):
                break
        status = ws["status"].view(-1, 2).cpu().numpy()
        return ws["x"].clone(), status[:, 0].copy(), status[:, 1].copy()

    def adjoint_solve(self, xk, rhs, use_disc=False, use_self=False, use_sdf=False, rtol=1e-9,
                      maxiter=5000, as_tensor=False, use_ipc=None):
        """One adjoint solve H_k lam = rhs on the free degrees of freedom, per-scene MINRES with the
        Jacobi preconditioner (reference implicit/adjoint.py:339-369); with affine bodies the
        reduced system P^T H P lam_r = P^T rhs, returning lam = P lam_r. Records minres_calls,
        minres_fail (scenes that did not converge), minres_last_info, minres_iters (the largest
        per-scene count), and minres_resid when minres_track_resid is set."""
        xw = self._t(xk)
        use = dict(use_disc=use_disc, use_sdf=use_sdf, use_self=use_self, use_ipc=use_ipc)
        if self._ipc_on(use_ipc):
            self._set_pairs(xw)
        rhs_t = self._t(rhs)
        diag = self._diag_precond(xw, **use)
        if self.affine is not None:
            b = self._reduce(rhs_t, torch.empty_like(rhs_t)) * self._fmask_red
            dred = self._reduce(diag, torch.empty_like(diag), square=True)
            dinv = self._fmask_red / torch.clamp(dred, min=1e-30)
        else:
            b = rhs_t * self._free
            dinv = self._free / diag
        lam, status, its = self._minres(xw, b, dinv, float(rtol), int(maxiter), **use)
        if self.affine is not None:
            lam = self._expand(lam, torch.empty_like(lam))
        self.minres_calls = getattr(self, "minres_calls", 0) + 1
        self.minres_iters = int(its.max(initial=0))
        bad = status != 1
        if bad.any():
            self.minres_fail = getattr(self, "minres_fail", 0) + int(bad.sum())
            self.minres_last_info = int(status[bad][0])
        if getattr(self, "minres_track_resid", False):
            if self.affine is not None:       # P^T H P lam_r - b, with P lam_r = lam
                kf = torch.zeros_like(lam)
                self._kv(xw, lam, kf, **use)
                r = self._reduce(kf, torch.empty_like(kf), vf=lam) + self._workspace()["mass"][:, None] * lam
                r = r * self._fmask_red - b
            else:
                r = self._hvp(xw, lam, **use) * self._free - b
            rel = float(torch.linalg.vector_norm(r)) / max(float(torch.linalg.vector_norm(b)), 1e-300)
            self.minres_resid = max(getattr(self, "minres_resid", 0.0), rel)
        return lam if as_tensor else lam.cpu().numpy()

    def adjoint_cg(self, xstar, b, diagH, use_self=False, include_mass=True, iters=8000, tol=1e-9):
        """Jacobi-preconditioned CG on the free degrees of freedom (for a positive definite H);
        records negative curvature it meets in _cg_info (reference :258-277)."""
        xw = self._t(xstar)
        free = self._free
        pre = free / torch.clamp(self._t(diagH), min=1e-30)
        lam = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        r = self._t(b) * free
        z = pre * r
        p = z.clone()
        rz = float((r * z).sum())
        r0 = float(torch.linalg.vector_norm(r))
        neg, it = 0, 0
        for it in range(1, iters + 1):
            Ap = self._hvp(xw, p * free, include_mass=include_mass, use_self=use_self) * free
            pAp = float((p * Ap).sum())
            if pAp <= 0.0:
                neg += 1
            a = rz / pAp
            lam = lam + a * p
            r = r - a * Ap
            if float(torch.linalg.vector_norm(r)) <= tol * r0 + 1e-30:
                break
            z = pre * r
            rz2 = float((r * z).sum())
            p = z + (rz2 / rz) * p
            rz = rz2
        self._cg_info = dict(iters=it, rel_resid=float(torch.linalg.vector_norm(r)) / (r0 + 1e-30),
                             neg_curv=neg)
        return lam.cpu().numpy()

    # ---- dJ/drho of one adjoint vector (reference :165-178, 280-285)
    def _drho_dev(self, xw, lamw):
        L = _lib.lib()
        s = self.s
        self._drho.zero_()
        _lib.check(L.d2f_elastic_drho_f64(self._p(xw), self._p(lamw), self._p(s.tri), self._p(s.Bm),
                                          self._p(s.area), self._p(self.mu0), self._p(self.lam0),
                                          self.n_tri, self.h2, self.Emin, self._p(self._drho),
                                          self._stream()), "d2f_elastic_drho_f64")
        return self._drho

    def drho(self, xstar, lam):
        return self._drho_dev(self._t(xstar), self._t(lam)).cpu().numpy().copy()

    # ---- objectives on the elastic force (reference :287-326)
    def elastic_force(self, x):
        """Elastic force on every node at positions x, (N, 2), with the current density."""
        L = _lib.lib()
        s = self.s
        xw = self._t(x)
        g = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        d = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        _lib.check(L.d2f_elastic_grad_diag_f64(self._p(xw), self._p(s.tri), self._p(s.Bm), self._p(s.area),
                                               self._p(s.mu), self._p(s.lam), self.n_tri, 1.0, 1,
                                               C.byref(self._tiles), self._p(g), self._p(d), self._stream()),
                   "elastic grad/diag")
        return -g.cpu().numpy()

    def force_objective(self, x, nodes, direction):
        """J = direction . (elastic force summed over `nodes`) at x; returns (J, dJ/dx, explicit
        dJ/drho). Use transient_adjoint(..., rtol=1e-14) with the dJ/dx seed, then add the
        explicit part."""
        nodes = np.asarray(nodes, np.int64)
        d = np.asarray(direction, np.float64)
        w = np.zeros((self.N, 2))
        w[nodes] = d
        J = float((self.elastic_force(x)[nodes] * d).sum())
        dJdx = -self.apply_H(x, w, include_mass=False, use_ipc=False) / self.h2
        return J, dJdx, self.drho(x, w / self.h2)

    # ---- the trajectory (reference :373-399)
    def _check_terms(self):
        s = self.s
        missing = []
        ipc_fric = s.ipc is not None and s.friction and (
            s.ipc["mu"] > 0.0 or (s.mu_node is not None and bool((s.mu_node > 0.0).any())))
        if s.tendon is not None:
            missing.append("Tendon")
        if (s.node is not None and s.node["fric"]) or ipc_fric:
            missing.append("BodyFriction")
        if s.springs is not None:
            missing.append("Springs")
        if s.activation is not None:
            missing.append("ActiveStrain")
        if missing:
            raise NotImplementedError(
                f"the adjoint operator has no tangent for {', '.join(sorted(missing))}: the gradient "
                f"would be of a different system. Differentiate without these actuators."
                + (" Friction between bodies has no tangent: set model.body_friction = 0 or "
                   "ImplicitIntegrator(friction=False)." if "BodyFriction" in missing else ""))

    # ---- dJ/dp of one adjoint vector (pressure chambers)
    def _dp_dev(self, xw, lamw):
        """-lam . dR/dp per chamber, in the model's chamber order."""
        L = _lib.lib()
        s = self.s
        out = torch.zeros(s.pressure["n_ch"], dtype=F64, device=self.dev)
        _lib.check(L.d2f_pressure_dp_f64(self._p(xw), self._p(lamw), self._p(s.pe_a), self._p(s.pe_b),
                                         self._p(s.pe_ch), s.pressure["n_pe"], self.h2, self._p(out),
                                         self._stream()), "d2f_pressure_dp_f64")
        model_order = torch.empty_like(out)
        model_order[torch.as_tensor(s.ch_order, device=self.dev)] = out
        return model_order

    def adjoint_step(self, xk, lam1, lam2, seed_k=None, drag=None, use_disc=False, use_self=False,
                     use_sdf=False, **minres_kw):
        """One step of the damped backward-Euler recursion on device tensors: solves
        H_k lam_k = seed_k + (1 + drag) M lam_{k+1} - drag M lam_{k+2} and returns (lam_k,
        dJ/drho contribution, dJ/dp contribution or None)."""
        dr = float(self.s.drag) if drag is None else float(drag)
        rhs = (1.0 + dr) * self._mass2 * lam1 - dr * self._mass2 * lam2
        if seed_k is not None:
            rhs = rhs + self._t(seed_k)
        xw = self._t(xk)
        lam = self.adjoint_solve(xw, rhs, use_disc=use_disc, use_self=use_self, use_sdf=use_sdf,
                                 as_tensor=True, **minres_kw)
        g = self._drho_dev(xw, lam).clone()
        gp = self._dp_dev(xw, lam) if self.s.pressure is not None else None
        return lam, g, gp

    def transient_adjoint(self, xs, seed, drag=None, use_disc=False, use_self=False, use_sdf=False,
                          pressures=None, **minres_kw):
        """xs = [x_0 .. x_T] checkpointed states; seed = {k: dJ/dx_k (N, 2)}. Returns dJ/drho
        (n_tri,) accumulated over the trajectory. With pressure chambers, dJ/dp per step and
        chamber (model order) lands in self.dJdp_steps (T, n_chambers); `pressures` (T,
        n_chambers) gives the pressures each step was taken with (default: the current ones)."""
        self._check_terms()
        T = len(xs) - 1
        self.refresh_mass()
        zero = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        lam = {T + 1: zero, T + 2: zero}
        g = torch.zeros(self.n_tri, dtype=F64, device=self.dev)
        s = self.s
        keep_p = None if s.pressure is None else s.ch_p.clone()
        self.dJdp_steps = None if s.pressure is None else np.zeros((T, s.pressure["n_ch"]))
        try:
            for k in range(T, 0, -1):
                if pressures is not None and s.pressure is not None:
                    s.set_pressure(np.asarray(pressures, np.float64)[k - 1])
                lam[k], gk, gp = self.adjoint_step(xs[k], lam[k + 1], lam[k + 2], seed.get(k), drag,
                                                   use_disc, use_self, use_sdf, **minres_kw)
                g += gk
                if gp is not None:
                    self.dJdp_steps[k - 1] = gp.cpu().numpy()
                lam.pop(k + 2, None)
        finally:
            if keep_p is not None:
                s.ch_p.copy_(keep_p)
        return g.cpu().numpy()
