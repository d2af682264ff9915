"""The oracle's implicit step: state, iteration loop and integrator glue
(reference implicit/solver.py:96-612, implicit/kernels.py:42-441, implicit/integrator.py:34-311).

    integ = OracleIntegrator(**same kwargs as ImplicitIntegrator)
    integ.simulate(model, dt)          # model: a finalized diff_2d_fem_amd.Model (host data)
    x = integ.solver.x

Test infrastructure only: the product never imports this.
"""
from __future__ import annotations

import math

import numpy as np

from . import terms as K
from .bodies import ActiveStrain, AffineBodies, ExternalForce, Pressure, RigidModes, Springs, Tendon
from .ipc import IPCContact
from .node import BodyFriction, NodeContact


class OracleSolver:
    """State and iteration of one batched implicit step, float64 on the CPU."""

    def __init__(self, me, cfg, driven, colliders=(), dt=1.0e-3, iter_max=150, eps=1e-4,
                 tol_x=0.0, conv_check=0, near_factor=4.0, alpha_max=3.0, area_shrink=0.1,
                 ccd_eta=0.9, friction=False, line_search=False, n_backtrack=4, ls_tol=1e-7):
        self.me, self.cfg = me, cfg
        self.dt = float(dt)
        self.iter_max, self.eps, self.tol_x = int(iter_max), float(eps), float(tol_x)
        self.conv_check = int(conv_check)
        self.near_factor, self.alpha_max = float(near_factor), float(alpha_max)
        self.area_shrink, self.ccd_eta = float(area_shrink), float(ccd_eta)
        self.friction = bool(friction)
        self.line_search, self.n_bt, self.ls_tol = bool(line_search), int(n_backtrack), float(ls_tol)
        self.terms, self.modes = [], []
        self.ipc = None
        self.node = None                 # NodeContact when body_contact is "node"
        self.affine = None
        self.cg_guard = 0.0
        self.B = int(me.n_envs)
        self.N = me.rest_pos.shape[0]
        self.n_tri = len(me.rest_area)
        self.n_b = len(me.boundary_nodes)
        self.x = np.asarray(me.rest_pos, np.float64).copy()
        z2 = lambda: np.zeros((self.N, 2))
        self.v, self.xn, self.xhat = z2(), z2(), z2()
        self.grad, self.grad_prev, self.p, self.diagH = z2(), z2(), z2(), z2()
        self.mass = np.asarray(me.mass, np.float64).copy()
        self.tri = np.asarray(me.tri, np.int64)
        self.Bm = np.asarray(me.D_m_inv, np.float64).copy()
        self.area = np.asarray(me.rest_area, np.float64).copy()
        self.mu = np.asarray(me.mu, np.float64).copy()
        self.lam = np.asarray(me.lam, np.float64).copy()
        self.bnd = np.asarray(me.boundary_nodes, np.int64)
        self.b2e = np.asarray(me.boundary2env, np.int64)
        self.node2env = np.asarray(me.node2env, np.int64)
        self.tri2env = np.asarray(me.tri2env, np.int64)
        self.free = np.ones((self.N, 2))
        self.bc_idx = np.asarray(driven, np.int64)
        self.free[self.bc_idx] = 0.0
        self.bc_target = np.zeros((len(self.bc_idx), 2))
        self.n_bc = len(self.bc_idx)
        zb = lambda: np.zeros(self.B)
        self.gPg, self.g_p, self.g_Py, self.y_p, self.y_Py = zb(), zb(), zb(), zb(), zb()
        self.beta, self.gTp, self.pHp = zb(), zb(), zb()
        self.alpha, self.p_inf, self.clr_min = zb(), zb(), zb()
        self.dE, self.dE0, self.Ener = zb(), zb(), zb()
        self.active = np.zeros(self.B, np.int32)
        self.iters = np.zeros(self.B, np.int32)
        self.dhat = float(cfg.ipc_d_hat)
        self.kappa = float(cfg.obstacle_stiffness)
        self.h2 = self.dt * self.dt
        self.inv_dt = 1.0 / self.dt
        self.drag = math.exp(-self.dt * cfg.damping)
        self.epsv = float(cfg.obstacle_friction_epsv)
        self.colliders = [dict(c) for c in colliders]
        self.iter_log = []

    def set_dt(self, dt):
        self.dt = float(dt)
        self.h2 = self.dt * self.dt
        self.inv_dt = 1.0 / self.dt
        self.drag = math.exp(-self.dt * self.cfg.damping)

    # ------------------------------------------------------------------ one iteration's parts
    def grad_diag(self):
        """Gradient and Jacobi diagonal of every term (reference implicit/solver.py:260-306)."""
        self.grad_prev = self.grad.copy()
        self.grad = self.free * (self.mass[:, None] * (self.x - self.xhat))
        self.diagH = np.repeat(self.mass[:, None], 2, 1)
        K.elastic_grad_diag(self.x, self.tri, self.Bm, self.area, self.mu, self.lam, self.h2, 1,
                            self.grad, self.diagH)
        for mode in self.modes:
            mode.after_elastic(self)
        for col in self.colliders:
            K.collider_grad_diag(col, self.x, self.bnd, self.dhat, self.kappa, self.h2, 1,
                                 self.grad, self.diagH)
        if self.node is not None:
            self.node.grad_diag(self.x, self.h2, 1, self.grad, self.diagH)
        if self.friction:
            for col in self.colliders:
                K.friction_grad_diag(col, self.x, self.xn, self.bnd, self.dhat, self.kappa, col["mu"],
                                     self.epsv, self.inv_dt, self.h2, self.free, self.grad, self.diagH)
        for term in self.terms:
            term.grad_diag(self)
        for mode in self.modes:
            mode.gather(self)
        self.grad *= self.free

    def direction(self, it):
        """Dai-Kou reductions, beta, restart guard and the new direction
        (reference implicit/kernels.py:168-230, implicit/solver.py:77-93, 308-328)."""
        Px = self.free[:, 0] / np.maximum(self.diagH[:, 0], 1.0e-12)
        Py = self.free[:, 1] / np.maximum(self.diagH[:, 1], 1.0e-12)
        g, pp = self.grad, self.p
        yy = g - self.grad_prev
        Pg0, Pg1 = Px * g[:, 0], Py * g[:, 1]
        Py0, Py1 = Px * yy[:, 0], Py * yy[:, 1]
        e = self.node2env
        for arr in (self.gPg, self.g_p, self.g_Py, self.y_p, self.y_Py):
            arr[:] = 0.0
        np.add.at(self.gPg, e, g[:, 0] * Pg0 + g[:, 1] * Pg1)
        np.add.at(self.g_p, e, g[:, 0] * pp[:, 0] + g[:, 1] * pp[:, 1])
        np.add.at(self.g_Py, e, g[:, 0] * Py0 + g[:, 1] * Py1)
        np.add.at(self.y_p, e, yy[:, 0] * pp[:, 0] + yy[:, 1] * pp[:, 1])
        np.add.at(self.y_Py, e, yy[:, 0] * Py0 + yy[:, 1] * Py1)
        for mode in self.modes:
            mode.reduce(self)
        b = np.zeros(self.B)
        if it > 0:
            ytp = self.y_p
            ok = np.abs(ytp) > 1.0e-30
            with np.errstate(divide="ignore", invalid="ignore"):
                bb = self.g_Py / ytp - self.y_Py * self.g_p / (ytp * ytp)
            b = np.where(ok, bb, 0.0)
        t = -self.gPg + b * self.g_p
        restart = t >= 0.0
        self.beta = np.where(restart, 0.0, b)
        self.gTp = np.where(restart, -self.gPg, t)
        if self.cg_guard > 0.0 and it > 0:
            hold = self.ipc.q_inf if (self.ipc is not None and self.ipc.pending) else np.zeros(self.B)
            g_ = (hold <= 0.0) & (self.y_p < self.cg_guard * self.alpha * self.pHp)
            self.beta = np.where(g_, 0.0, self.beta)
            self.gTp = np.where(g_, -self.gPg, self.gTp)
        bb = self.beta[e]
        px = -self.free[:, 0] / np.maximum(self.diagH[:, 0], 1.0e-12) * g[:, 0] + bb * self.p[:, 0]
        py = -self.free[:, 1] / np.maximum(self.diagH[:, 1], 1.0e-12) * g[:, 1] + bb * self.p[:, 1]
        self.p = np.stack([px, py], 1) * self.free
        for mode in self.modes:
            mode.direction(self)

    def curvature(self):
        """p^T H p, step norms and clearances (reference implicit/solver.py:330-378)."""
        self.pHp[:] = 0.0
        self.alpha[:] = self.alpha_max
        self.p_inf[:] = 0.0
        self.clr_min[:] = K.BIG
        e = self.node2env
        np.add.at(self.pHp, e, self.mass * (self.p * self.p).sum(1))
        np.maximum.at(self.p_inf, e, np.linalg.norm(self.p, axis=1))
        K.elastic_pHp(self.x, self.p, self.tri, self.Bm, self.area, self.mu, self.lam, self.h2, 1,
                      self.tri2env, self.pHp)
        for col in self.colliders:
            K.collider_pHp(col, self.x, self.p, self.bnd, self.b2e, self.dhat, self.kappa, self.h2, 1,
                           self.pHp, self.clr_min)
        if self.friction:
            for col in self.colliders:
                K.friction_pHp(col, self.x, self.xn, self.p, self.bnd, self.b2e, self.dhat, self.kappa,
                               col["mu"], self.epsv, self.inv_dt, self.h2, self.free, self.pHp)
        if self.node is not None:
            self.node.pHp(self.x, self.p, self.h2, 1, self.pHp)
        for term in self.terms:
            term.pHp(self)

    def caps(self, it):
        """Inversion and collider step caps, then the per-env step length and convergence
        tests (reference implicit/kernels.py:304-423)."""
        K.inversion_alpha(self.x, self.p, self.tri, self.tri2env, self.area_shrink, self.alpha)
        for col in self.colliders:
            K.collider_ccd(col, self.x, self.p, self.bnd, self.b2e, self.ccd_eta, self.alpha)
        for e in range(self.B):
            if self.active[e] == 0:
                continue
            ph = self.pHp[e]
            a = self.alpha[e]
            if ph > 0.0:
                a = min(a, -self.gTp[e] / ph)
            if self.clr_min[e] < self.near_factor * self.dhat and self.p_inf[e] > 0.0:
                a = min(a, 0.5 * self.dhat / self.p_inf[e])
            if a <= 0.0:
                self.active[e] = 0
                self.alpha[e] = 0.0
                continue
            self.alpha[e] = a
            if self.tol_x > 0.0 and a * self.p_inf[e] < self.tol_x:
                self.active[e] = 0
                self.iters[e] += 1
                continue
            d = -a * self.gTp[e] - 0.5 * a * a * ph
            self.dE[e] = d
            if it == 0:
                self.dE0[e] = max(d, 1.0e-300)
            elif d < self.eps * self.dE0[e]:
                self.active[e] = 0
            self.iters[e] += 1

    def update(self):
        act = (self.active[self.node2env] != 0)
        self.x[act] = self.x[act] + self.alpha[self.node2env[act]][:, None] * self.p[act]
        for mode in self.modes:
            mode.advance(self)

    def energy_into(self, buf):
        """Total incremental potential per env (reference implicit/solver.py:453-498)."""
        buf[:] = 0.0
        d = self.x - self.xhat
        np.add.at(buf, self.node2env, 0.5 * self.mass * (d * d).sum(1))
        K.elastic_energy(self.x, self.tri, self.Bm, self.area, self.mu, self.lam, self.h2,
                         self.tri2env, buf)
        for col in self.colliders:
            K.collider_energy(col, self.x, self.bnd, self.b2e, self.dhat, self.kappa, self.h2, buf)
        if self.node is not None:
            self.node.energy(self.x, self.h2, buf)
        if self.friction:
            for col in self.colliders:
                K.friction_energy(col, self.x, self.xn, self.bnd, self.b2e, self.dhat, self.kappa,
                                  col["mu"], self.epsv, self.inv_dt, self.h2, self.free, buf)
        for term in self.terms:
            term.energy(self, buf)

    def line_search_update(self):
        """Energy backtracking gate (reference implicit/solver.py:403-425)."""
        Eprev = np.zeros(self.B)
        self.energy_into(Eprev)
        x0 = self.x.copy()
        for mode in self.modes:
            mode.save(self)
        act = self.active[self.node2env] != 0
        for _ in range(self.n_bt):
            self.x[act] = x0[act] + self.alpha[self.node2env[act]][:, None] * self.p[act]
            for mode in self.modes:
                mode.trial(self)
            self.energy_into(self.Ener)
            up = (self.active != 0) & (self.Ener > Eprev + self.ls_tol * (np.abs(Eprev) + 1.0))
            self.alpha[up] *= 0.5
        self.x[act] = x0[act] + self.alpha[self.node2env[act]][:, None] * self.p[act]
        for mode in self.modes:
            mode.trial(self)

    # ------------------------------------------------------------------ the loop
    def iterate(self):
        for it in range(self.iter_max):
            if self.ipc is not None:
                self.ipc.before_iteration(self, it)
            if self.conv_check and it and it % self.conv_check == 0:
                if int(self.active.sum()) == 0:
                    break
            self.grad_diag()
            self.direction(it)
            self.curvature()
            self.caps(it)
            if not self.line_search:
                self.update()
            else:
                self.line_search_update()

    def step_targets(self):
        """One implicit step (reference implicit/solver.py:433-451)."""
        self.xn = self.x.copy()
        self.xhat = self.x + self.dt * self.drag * self.v
        if self.ipc is None:
            self.x[self.bc_idx] = self.bc_target
            self.xhat[self.bc_idx] = self.bc_target
        else:
            self.ipc.advance_bc(self)
        self.dE0[:] = 0.0
        self.active[:] = 1
        self.iters[:] = 0
        self.p[:] = 0.0
        self.grad[:] = 0.0
        for mode in self.modes:
            mode.begin_step(self)
        self.iterate()
        self.v = (self.x - self.xn) * (1.0 / self.dt)


class OracleIntegrator:
    """The reference ImplicitIntegrator's build and actuator glue over OracleSolver
    (reference implicit/integrator.py:34-311)."""

    def __init__(self, max_iter=None, eps=None, self_collision=True, friction=True,
                 near_factor=4.0, alpha_max=3.0, area_shrink=0.1, ccd_eta=0.9, self_radius=None,
                 tol_x=0.0, conv_check=8, line_search=False, n_backtrack=4, ls_tol=1e-7,
                 rigid_modes=True, iters_per_cfl=2.0, max_iter_cap=2000, cg_restart=None):
        self.self_radius = self_radius
        self.max_iter = None if max_iter is None else int(max_iter)
        self.iters_per_cfl, self.max_iter_cap = float(iters_per_cfl), int(max_iter_cap)
        self.eps = float(1e-10 if eps is None else eps)
        self.tol_x, self.conv_check = float(tol_x), int(conv_check)
        self.line_search, self.n_backtrack, self.ls_tol = bool(line_search), int(n_backtrack), float(ls_tol)
        self.self_collision, self.friction = bool(self_collision), bool(friction)
        self.near_factor, self.alpha_max = float(near_factor), float(alpha_max)
        self.area_shrink, self.ccd_eta = float(area_shrink), float(ccd_eta)
        self.rigid_modes = bool(rigid_modes)
        self.cg_restart = cg_restart
        self.solver = None
        self._act_version = None
        self._ctrl_version = None

    def cfl_number(self, model, dt):
        me = model.me
        mu, lam = np.asarray(me.mu, np.float64), np.asarray(me.lam, np.float64)
        soft = np.ones(len(mu), bool)
        for a in model._affine:
            soft[a["elems"]] = False
        C = 0.0
        for h, b in enumerate(model._bodies):
            el = model.body_elements(h)
            el = el[soft[el]]
            if len(el) and b.role != "kinematic":
                cp = np.sqrt((lam[el] + 2.0 * mu[el]).max() / b.rho)
                C = max(C, cp * dt / b.cfl_length())
        return float(C)

    def iters_for(self, model, dt):
        if self.max_iter is not None:
            return self.max_iter
        n = int(np.ceil(self.iters_per_cfl * self.cfl_number(model, dt)))
        lo = 120 if self.self_collision and model.body_contact == "ipc" else 40
        return int(np.clip(n, lo, self.max_iter_cap))

    def _build(self, model, dt):
        s = OracleSolver(model.me, model.cfg, model.all_constrained(), colliders=model.collider_records(),
                         dt=dt, iter_max=self.iters_for(model, dt), eps=self.eps, tol_x=self.tol_x,
                         conv_check=self.conv_check, line_search=self.line_search,
                         n_backtrack=self.n_backtrack, ls_tol=self.ls_tol,
                         near_factor=self.near_factor, alpha_max=self.alpha_max,
                         area_shrink=self.area_shrink, ccd_eta=self.ccd_eta, friction=self.friction)
        if model._v0 is not None and np.any(model._v0):
            s.v = np.asarray(model._v0, np.float64).copy()
        s.affine = AffineBodies(s, model._affine) if model._affine else None
        if s.affine is not None:
            s.modes.append(s.affine)
            s.terms.append(s.affine)
        mu_bb = model.body_friction if self.friction else 0.0
        mu_node = model._body_mu_per_node() if self.friction else None
        if model.body_contact not in ("ipc", "node"):
            raise ValueError(f"model.body_contact must be 'ipc' or 'node', not {model.body_contact!r}")
        if self.self_collision and model.body_contact == "node":
            r = self.self_radius if self.self_radius is not None else 1.5 * model.dx
            s.node = NodeContact(s, model.cfg.contact_stiffness, r)
            if mu_bb > 0.0 or (mu_node is not None and np.max(mu_node) > 0.0):
                s.terms.append(BodyFriction(s, s.node, mu_bb, mu_node))
        if self.self_collision and model.body_contact == "ipc":
            s.ipc = IPCContact(s, mu_bb, dhat=model.dhat, kappa=model.body_kappa,
                               reach=max(2.0 * model.dx, 2.0 * model.dhat), mu_node=mu_node)
            s.terms.append(s.ipc)
            gap = s.ipc.min_distance(s)
            if not gap.min() > 0.0:
                raise ValueError(f"two surfaces touch in environment {int(gap.argmin())} before the "
                                 f"first step")
        s.cg_guard = float((0.5 if s.ipc is not None else 0.0) if self.cg_restart is None
                           else self.cg_restart)
        if self.rigid_modes:
            rm = RigidModes.build(s, skip_nodes=(s.affine.slaved if s.affine is not None else None))
            if rm is not None:
                s.modes.append(rm)
        self.solver = s
        self._force = None

    def _sync_actuation(self, model):
        s = self.solver
        act = model.actuation
        if self._act_version != act.structure:
            keep = [t for t in s.terms if isinstance(t, ExternalForce) or getattr(t, "persistent", False)]
            s.terms = list(keep)
            self._force = next((t for t in keep if isinstance(t, ExternalForce)), None)
            self._pressure = self._tendon = self._spring = self._strain = None
            if act.springs:
                lay = act.spring_layout()
                self._spring = Springs(s, lay["a"], lay["b"], lay["stiffness"], lay["env"])
                s.terms.append(self._spring)
            if act.chambers:
                lay = act.pressure_layout(model.body_env)
                self._pressure = Pressure(s, lay["ea"], lay["eb"], lay["chamber"], lay["env"], lay["ref"])
                s.terms.append(self._pressure)
            if act.tendons:
                lay = act.tendon_layout(model.body_env)
                node, prv, nxt, tid = [], [], [], []
                for i, path in enumerate(lay["paths"]):
                    o, n = len(node), len(path)
                    node += list(path)
                    tid += [i] * n
                    prv += [-1] + list(range(o, o + n - 1))
                    nxt += list(range(o + 1, o + n)) + [-1]
                self._tendon = Tendon(s, node, prv, nxt, tid, lay["env"], lay["stiffness"], lay["length"])
                s.terms.append(self._tendon)
            if act.regions:
                lay = act.region_layout()
                self._strain = ActiveStrain(s, lay["elems"], lay["region"], lay["n"])
            self._act_version = act.structure
            self._ctrl_version = None
        if self._ctrl_version != act.commands:
            if self._pressure is not None:
                self._pressure.set([c.pressure for c in act.chambers])
            if self._tendon is not None:
                self._tendon.set([t.tension for t in act.tendons], [t.length for t in act.tendons])
            if self._strain is not None:
                self._strain.set(s, np.stack([r.A for r in act.regions]))
            if self._spring is not None:
                self._spring.set([k.anchor for k in act.springs], [k.length for k in act.springs])
            f = act.external_force(s.mass)
            if f is not None:
                if self._force is None:
                    self._force = ExternalForce(s)
                    s.terms.insert(0, self._force)
                self._force.set(f)
            self._ctrl_version = act.commands

    def attach(self, model, dt):
        if self.solver is None:
            self._build(model, dt)
        s = self.solver
        if abs(s.dt - dt) > 1e-12 * dt:
            s.set_dt(dt)
            s.iter_max = self.iters_for(model, dt)
        self._sync_actuation(model)
        return s

    def simulate(self, model, dt):
        s = self.attach(model, dt)
        s.bc_target = np.asarray(model._targets, np.float64)[s.bc_idx]
        s.step_targets()

    def stats(self):
        s = self.solver
        out = dict(iters=s.iters.copy(), converged=(s.active == 0), max_iter=s.iter_max)
        if s.ipc is not None and s.n_bc:
            out["bc_lag"] = float(np.linalg.norm(s.bc_target - s.x[s.bc_idx], axis=1).max())
        return out
