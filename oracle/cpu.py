"""The oracle's step in compiled C++ (oracle/cpu_step.cpp) — TEST AND BASELINE INFRASTRUCTURE
ONLY: the product never imports it.

    cs = CpuStepper(OracleIntegrator(...), model, dt, threads=1)
    cs.simulate(model, dt)             # one implicit step of every environment
    x = cs.solver.x

CpuStepper attaches an OracleIntegrator to the model (the numpy oracle's build: the same
terms, parameters and initial state) and then steps the oracle solver's own arrays with
d2f_cpu_step instead of OracleSolver.step_targets: serially (threads=1, the reference's CPU
semantics of one kernel thread after another) or with OpenMP threads over environments. It
covers the contact-heavy configuration's terms (elasticity, half-space and disc colliders
with friction, IPC contact between bodies with friction and the kinematic sweep, affine
bodies, an external force) and refuses models with others. tests/test_oracle_cpu.py checks it
against the numpy oracle.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libd2f_cpu_oracle.so"
P = C.c_void_p


class _Scene(C.Structure):
    _fields_ = ([(k, C.c_int64) for k in ("B", "N", "T", "n_b", "n_e", "n_bc", "n_ag", "n_col")]
                + [(k, P) for k in ("x", "v", "xn", "xhat", "grad", "grad_prev", "p", "diag", "bc_q",
                                    "mass", "free_", "fext", "node_start", "tri", "Bm", "area", "mu", "lam",
                                    "tri_start", "bnd", "bnd_start", "ea", "eb", "ela", "elb", "e_start", "ca",
                                    "cb", "excl_ptr", "excl_idx", "mu_node", "bc_idx", "bc_start", "bc_target",
                                    "an_node", "an_start", "ag_start", "an_w", "ag_Mu", "ag_La", "ag_q", "col",
                                    "iters", "active")])


class _Params(C.Structure):
    _fields_ = ([(k, C.c_double) for k in ("dt", "drag", "dhat", "kappa", "epsv", "eps", "tol_x", "near_factor",
                                           "alpha_max", "area_shrink", "ccd_eta", "cg_guard", "ipc_mu", "ipc_dhat",
                                           "ipc_kappa", "ipc_reach", "ipc_gap")]
                + [(k, C.c_int32) for k in ("iter_max", "conv_check", "friction", "ipc_on", "ipc_per_body", "pad_")])


def build(force=False):
    """Compile cpu_step.cpp into libd2f_cpu_oracle.so beside it (g++, OpenMP, portable code)."""
    import subprocess
    src = HERE / "cpu_step.cpp"
    if LIB.exists() and not force and LIB.stat().st_mtime >= src.stat().st_mtime:
        return LIB
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
           str(src), "-o", str(LIB)]
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: build it with oracle.cpu.build()")
        L = C.CDLL(str(LIB))
        L.d2f_cpu_step.argtypes = [C.POINTER(_Scene), C.POINTER(_Params), C.c_int]
        L.d2f_cpu_step.restype = C.c_int
        L.d2f_cpu_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class CpuStepper:
    def __init__(self, integrator, model, dt, threads=1):
        from .bodies import AffineBodies, ExternalForce
        from .ipc import IPCContact
        self.oi = integrator
        s = integrator.attach(model, dt)
        self.solver = s
        self.threads = int(threads)
        for t in s.terms:
            if not isinstance(t, (ExternalForce, AffineBodies, IPCContact)):
                raise NotImplementedError(f"the compiled step has no {type(t).__name__}")
        if s.node is not None or s.line_search or any(not isinstance(m, AffineBodies) for m in s.modes):
            raise NotImplementedError("the compiled step has no node contact, line search or rigid modes")
        if any(int(c["kind"]) not in (0, 4) for c in s.colliders):
            raise NotImplementedError("the compiled step has disc and half-space colliders only")
        me = model.me
        keep = self._keep = []

        def arr(a, dtype):
            a = _c(a, dtype)
            keep.append(a)
            return a.ctypes.data_as(P)
        for k in ("x", "v", "xn", "xhat", "grad", "grad_prev", "p", "diagH"):
            setattr(s, k, _c(getattr(s, k), np.float64))
        self.bc_q = np.zeros((s.N, 2))
        sc = self.sc = _Scene()
        sc.B, sc.N, sc.T, sc.n_b = s.B, s.N, s.n_tri, s.n_b
        B = s.B
        starts = lambda keys, n: np.searchsorted(np.asarray(keys), np.arange(n + 1)).astype(np.int64)
        sc.node_start = arr(starts(s.node2env, B), np.int64)
        sc.tri, sc.Bm = arr(s.tri, np.int64), arr(np.asarray(s.Bm).reshape(-1, 4), np.float64)
        sc.area = arr(s.area, np.float64)
        sc.tri_start = arr(starts(s.tri2env, B), np.int64)
        sc.mass, sc.free_ = arr(s.mass, np.float64), arr(s.free, np.float64)
        sc.bnd, sc.bnd_start = arr(s.bnd, np.int64), arr(np.asarray(me.boundary_starts), np.int64)
        sc.excl_ptr, sc.excl_idx = arr(me.excl_ptr, np.int64), arr(me.excl_idx, np.int32)
        env_bc = np.asarray(s.node2env)[s.bc_idx]
        if (np.diff(env_bc) < 0).any():
            raise ValueError("kinematic nodes are not grouped by environment")
        sc.n_bc, sc.bc_idx, sc.bc_start = len(s.bc_idx), arr(s.bc_idx, np.int64), arr(starts(env_bc, B), np.int64)
        ip = s.ipc
        if ip is not None:
            sc.n_e = len(ip.ea)
            if (np.diff(ip.e2env) < 0).any():
                raise ValueError("boundary edges are not grouped by environment")
            for k in ("ea", "eb", "ela", "elb"):
                setattr(sc, k, arr(getattr(ip, k), np.int64))
            sc.ca, sc.cb = arr(ip.ca, np.float64), arr(ip.cb, np.float64)
            sc.e_start = arr(starts(ip.e2env, B), np.int64)
            sc.mu_node = None if ip.mu_node is None else arr(ip.mu_node, np.float64)
        af = s.affine
        if af is not None:
            order = np.argsort(af.env, kind="stable")             # groups by environment
            self._gorder = order
            node, w, st = [], [], [0]
            for g in order:
                k = slice(af.first[g], af.first[g + 1])
                node.append(af.node[k])
                w.append(af.w[k])
                st.append(st[-1] + len(af.node[k]))
            sc.n_ag = af.n
            sc.an_node, sc.an_w = arr(np.concatenate(node), np.int64), arr(np.concatenate(w), np.float64)
            sc.an_start, sc.ag_start = arr(st, np.int64), arr(starts(af.env[order], B), np.int64)
            self.ag_q = _c(af.q[order], np.float64)
        cols = [[c["kind"], c["cx"], c["cy"], c["rx"], c["ry"], c["rref"], c["r"], c["mu"]] for c in s.colliders]
        sc.n_col = len(cols)
        sc.col = arr(np.asarray(cols, np.float64).reshape(-1), np.float64) if cols else None
        self.iters = np.zeros(B, np.int32)
        self.active = np.zeros(B, np.int32)

    def _params(self):
        s = self.solver
        q = _Params()
        for k in ("dt", "drag", "dhat", "kappa", "epsv", "eps", "tol_x", "near_factor", "alpha_max",
                  "area_shrink", "ccd_eta", "cg_guard"):
            setattr(q, k, float(getattr(s, k)))
        q.iter_max, q.conv_check, q.friction = s.iter_max, s.conv_check, int(s.friction)
        ip = s.ipc
        if ip is not None:
            q.ipc_on, q.ipc_per_body = 1, int(ip.per_body)
            q.ipc_mu, q.ipc_dhat, q.ipc_kappa, q.ipc_reach, q.ipc_gap = ip.mu, ip.dhat, ip.kappa, ip.reach, ip.gap
        return q

    def simulate(self, model, dt):
        from .bodies import ExternalForce
        s = self.oi.attach(model, dt)
        sc = self.sc
        s.bc_target = _c(np.asarray(model._targets, np.float64)[s.bc_idx], np.float64)
        f = next((t.f for t in s.terms if isinstance(t, ExternalForce)), None)
        self._fext = None if f is None else _c(f, np.float64)
        sc.fext = None if self._fext is None else self._fext.ctypes.data_as(P)
        self._mu, self._lam = _c(s.mu, np.float64), _c(s.lam, np.float64)
        sc.mu, sc.lam = self._mu.ctypes.data_as(P), self._lam.ctypes.data_as(P)
        sc.bc_target = s.bc_target.ctypes.data_as(P)
        for k, a in (("x", s.x), ("v", s.v), ("xn", s.xn), ("xhat", s.xhat), ("grad", s.grad),
                     ("grad_prev", s.grad_prev), ("p", s.p), ("diag", s.diagH), ("bc_q", self.bc_q)):
            setattr(sc, k, a.ctypes.data_as(P))
        if s.affine is not None:
            self._Mu = _c(s.affine.Mu[self._gorder], np.float64)
            self._La = _c(s.affine.La[self._gorder], np.float64)
            sc.ag_Mu, sc.ag_La = self._Mu.ctypes.data_as(P), self._La.ctypes.data_as(P)
            sc.ag_q = self.ag_q.ctypes.data_as(P)
        sc.iters, sc.active = self.iters.ctypes.data_as(P), self.active.ctypes.data_as(P)
        q = self._params()
        rc = lib().d2f_cpu_step(C.byref(sc), C.byref(q), self.threads)
        if rc != 0:
            raise RuntimeError("d2f_cpu_step refused its input")
        if s.affine is not None:
            s.affine.q[self._gorder] = self.ag_q
        s.iters[:] = self.iters
        s.active[:] = self.active
