"""Device state of one batched implicit solve, in PyTorch-ROCm tensors, and the calls into
the HIP library that advance it (reference implicit/solver.py:96-612, with the set-up of
implicit/affine.py, implicit/ipc_contact.py and implicit/actuation.py).

The host side builds index and constant arrays once (numpy, like the reference), then every
step is one call into libd2fem_hip.so; positions stay on the GPU and are read back only on
request (get_x).
"""
from __future__ import annotations

import ctypes as C
import math
import warnings

import numpy as np
import torch
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from . import _lib

I32, F64, U8, I64 = torch.int32, torch.float64, torch.uint8, torch.int64


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU: the implicit step runs only on the GPU (HIP); there is no CPU path")
    return torch.device("cuda")


class DeviceSolver:
    """Everything the kernels read and write, for every environment of one Model."""

    def __init__(self, me, cfg, driven, colliders=(), dt=1.0e-3, iter_max=150, eps=1e-4, tol_x=0.0,
                 conv_check=0, near_factor=4.0, alpha_max=3.0, area_shrink=0.1, ccd_eta=0.9,
                 friction=False, line_search=False, n_backtrack=4, ls_tol=1e-7, path="scene",
                 precision="double"):
        self.dev = _dev()
        self.me, self.cfg = me, cfg
        self.path, self.precision = path, precision
        self.iter_max, self.eps, self.tol_x = int(iter_max), float(eps), float(tol_x)
        self.conv_check = int(conv_check)
        self.near_factor, self.alpha_max = float(near_factor), float(alpha_max)
        self.area_shrink, self.ccd_eta = float(area_shrink), float(ccd_eta)
        self.friction = bool(friction)
        self.line_search, self.n_bt, self.ls_tol = bool(line_search), int(n_backtrack), float(ls_tol)
        self.cg_guard = 0.0
        self.ipc = None
        self.node = None                   # "node" contact between bodies, when set up
        self.affine = None
        self.rigid = None
        self.B = int(me.n_envs)
        self.N = me.rest_pos.shape[0]
        self.n_tri = len(me.rest_area)
        self.n_b = len(me.boundary_nodes)
        self.set_dt(dt)
        t = self._t
        z2 = lambda: torch.zeros((self.N, 2), dtype=F64, device=self.dev)
        self.x = t(me.rest_pos, F64)
        self.v, self.xn, self.xhat, self.x0 = z2(), z2(), z2(), z2()
        self.grad, self.grad_prev, self.p, self.diag, self.diag_el = z2(), z2(), z2(), z2(), z2()
        self.bc_q = z2()
        self.mass_np = np.asarray(me.mass, np.float64)
        self.mass = t(self.mass_np, F64)
        self.tri = t(np.asarray(me.tri, np.int32), I32)
        self.Bm = t(np.asarray(me.D_m_inv, np.float64).reshape(-1, 4), F64)
        self.area = t(me.rest_area, F64)
        self.mu_np = np.asarray(me.mu, np.float64).copy()
        self.lam_np = np.asarray(me.lam, np.float64).copy()
        self.mu = t(self.mu_np, F64)
        self.lam = t(self.lam_np, F64)
        self.node2env_np = np.asarray(me.node2env, np.int64)
        self.node2env = t(self.node2env_np, I32)
        self.block_size = 512              # scene path: 2 waves per SIMD, no register spills
        self.skin = self.SKIN              # scene path: pair-list margin (x the contact reach)
        self.lds_order = self.LDS_ORDER    # scene path: which optional LDS parts come first
        self._layout = None
        self._route_t = None
        self._order_t = None
        self.longest_first = True           # scene kernel: launch the largest scenes first
        self.n_cu = torch.cuda.get_device_properties(self.dev).multi_processor_count if torch.cuda.is_available() else 256
        # scene path: step through d2f_step_scene_noalloc (the library's static work arrays, no
        # per-call allocation; at most NOALLOC_MAX_NODES nodes, one solver per device at a time)
        self.noalloc = False
        self.excl_bits = self.excl_bits_off = None     # bit-packed exclusion (with contact)
        self.excl_words, self.excl_words_max = np.zeros(self.B, np.int64), 0
        self.scene_envs = np.arange(self.B)
        self.plain_envs = np.zeros(0, np.int64)
        self.tri2env = t(np.asarray(me.tri2env), I32)
        from .layout import element_tiles
        self.el_tiles = element_tiles(me.tri, self.N, device=self.dev)
        self.node_start = t(np.asarray(me.node_starts), I32)
        self.bnd = t(np.asarray(me.boundary_nodes), I32)
        self.b2e = t(np.asarray(me.boundary2env), I32)
        self.bnd_start = t(np.asarray(me.boundary_starts), I32)
        self.free_np = np.ones((self.N, 2))
        drv = np.asarray(driven, np.int64)
        # kinematic nodes grouped by environment (each one's target follows it)
        self.bc_idx_np = drv[np.argsort(self.node2env_np[drv], kind="stable")] if len(drv) else drv
        self.free_np[self.bc_idx_np] = 0.0
        self.free = t((self.free_np[:, 0] > 0.5).astype(np.uint8), U8)
        self.bc_idx = t(self.bc_idx_np, I32)
        self.n_bc = len(self.bc_idx_np)
        self.bc_start = t(np.searchsorted(self.node2env_np[self.bc_idx_np], np.arange(self.B + 1)), I32)
        self.bc_target = torch.zeros((max(self.n_bc, 1), 2), dtype=F64, device=self.dev)
        zb = lambda: torch.zeros(self.B, dtype=F64, device=self.dev)
        for k in ("gPg", "g_p", "g_Py", "y_p", "y_Py", "beta", "gTp", "pHp", "alpha", "p_inf",
                  "clr_min", "dE", "dE0", "Ener", "Eprev", "clr_ipc", "q_inf", "sweep_beta", "tc"):
            setattr(self, k, zb())
        self.active = torch.zeros(self.B, dtype=I32, device=self.dev)
        self.iters = torch.zeros(self.B, dtype=I32, device=self.dev)
        self.colliders = [dict(c) for c in colliders]
        if len(self.colliders) > _lib.MAX_COLLIDERS:
            raise ValueError(f"at most {_lib.MAX_COLLIDERS} colliders are supported")
        self.fext = None
        self.pressure = None
        self.tendon = None
        self.springs = None
        self.activation = None
        self._empty_i = torch.zeros(1, dtype=I32, device=self.dev)
        self._empty_f = torch.zeros(8, dtype=F64, device=self.dev)

    def _t(self, a, dtype):
        return torch.as_tensor(np.ascontiguousarray(a), device=self.dev).to(dtype).contiguous()

    def set_dt(self, dt):
        self.dt = float(dt)
        self.h2 = self.dt * self.dt
        self.inv_dt = 1.0 / self.dt
        self.drag = math.exp(-self.dt * self.cfg.damping)

    # ------------------------------------------------------------------ contact between bodies
    def setup_ipc(self, mu, dhat, kappa, reach, gap=0.1, mu_node=None):
        """IPC node-edge contact data (reference implicit/ipc_contact.py:353-386)."""
        me = self.me
        edges = np.asarray(me.boundary_edges, np.int64).reshape(-1, 2)
        env = self.node2env_np[edges[:, 0]] if len(edges) else np.zeros(0, np.int64)
        pos = np.full(self.N, -1, np.int64)
        bstart = np.asarray(me.boundary_starts)
        pos[np.asarray(me.boundary_nodes)] = np.arange(self.n_b) - bstart[np.asarray(me.boundary2env)]
        m = np.bincount(edges.reshape(-1), minlength=self.N).astype(np.float64)
        share = (m - 1.0) / np.maximum(m, 1.0)
        t = self._t
        self.ipc = dict(mu=float(mu), dhat=float(dhat), kappa=float(kappa), reach=float(reach),
                        gap=float(gap), per_body=mu_node is not None, pending=False)
        self.ea, self.eb = t(edges[:, 0], I32), t(edges[:, 1], I32)
        self.ela, self.elb = t(pos[edges[:, 0]], I32), t(pos[edges[:, 1]], I32)
        self.e2env = t(env, I32)
        self.ca, self.cb = t(share[edges[:, 0]], F64), t(share[edges[:, 1]], F64)
        self.mu_node = None if mu_node is None else t(mu_node, F64)
        self.n_e = len(edges)
        if len(env) and (np.diff(env) < 0).any():
            raise RuntimeError("boundary edges are not grouped by environment")
        self.e_start = t(np.searchsorted(env, np.arange(self.B + 1)), I32)
        self.max_nb = int(np.max(np.asarray(me.boundary_per_env))) if self.B else 0
        self._pack_exclusion()
        self._layout = None                 # the scene layout plans LDS for the contact lists

    def _pack_exclusion(self):
        """Each scene's exclusion, bit-packed from the mesh's lists, 32 per word (bit
        jp * n_b_e + j of scene e's words: boundary nodes jp and j do not contact): excl_bits,
        excl_bits_off, and the largest scene's word count excl_words_max. The device never
        holds the n_b_e x n_b_e byte matrices of the reference."""
        me = self.me
        nbe = np.asarray(me.boundary_per_env, np.int64)
        wcount = (nbe * nbe + 31) // 32
        woff = np.concatenate([[0], np.cumsum(wcount)]).astype(np.int64)
        ptr = np.asarray(me.excl_ptr, np.int64)
        idx = np.asarray(me.excl_idx, np.int64)
        bstart = np.asarray(me.boundary_starts, np.int64)
        n_b = len(ptr) - 1
        row = np.repeat(np.arange(n_b), np.diff(ptr))                    # global boundary index
        env = np.asarray(me.boundary2env, np.int64)[row]
        bit = woff[env] * 32 + (row - bstart[env]) * nbe[env] + idx         # global bit position
        words = np.zeros(int(woff[-1]), np.uint32)
        np.bitwise_or.at(words, bit // 32, (np.uint32(1) << (bit % 32).astype(np.uint32)))
        self.excl_bits = self._t(words.view(np.int32), I32)
        self.excl_bits_off = self._t(woff, I64)
        self.excl_words = wcount
        self.excl_words_max = int(wcount.max(initial=0))

    def setup_node(self, stiffness, radius, mu, mu_node=None, friction=False):
        """The "node" penalty contact between boundary nodes of one environment, with Coulomb
        friction on its pairs when `friction` (reference implicit/solver.py:215-226,
        implicit/body_contact.py:170-210)."""
        me = self.me
        t = self._t
        self.node = dict(k=float(stiffness), r=float(radius), mu=float(mu), fric=bool(friction),
                         per_body=mu_node is not None)
        self.mu_node = None if mu_node is None else t(mu_node, F64)
        self.max_nb = int(np.max(np.asarray(me.boundary_per_env))) if self.B else 0
        self._pack_exclusion()

    def refresh_exclusion(self):
        if self.ipc is not None or self.node is not None:
            self._pack_exclusion()
            self._layout = None

    # ------------------------------------------------------------------ rigid bodies
    def setup_affine(self, groups):
        """Rigid regions as affine bodies (reference implicit/affine.py:282-328)."""
        me = self.me
        tri = np.asarray(me.tri)
        X = np.asarray(me.rest_pos, np.float64)
        area = np.asarray(me.rest_area, np.float64)
        node, grp, w, Mu, La, env, gel = [], [], [], [], [], [], []
        owner = np.full(self.N, -1)
        genv = [int(np.asarray(me.tri2env)[np.asarray(g["elems"], np.int64)[0]]) for g in groups]
        order = np.argsort(genv, kind="stable")          # groups by environment
        self.ag_order = order
        for b, gi in enumerate(order):
            g = groups[gi]
            el = np.asarray(g["elems"], np.int64)
            nd = np.unique(tri[el].reshape(-1))
            if (owner[nd] >= 0).any():
                raise ValueError("two rigid regions share a node; merge them or leave a gap of at "
                                 "least one soft element between them")
            if (self.free_np[nd] == 0).any():
                raise ValueError("a rigid region contains kinematically constrained nodes. A body "
                                 "whose motion is PRESCRIBED needs no rigid region")
            owner[nd] = b
            gel.append(el)
            c = (self.mass_np[nd, None] * X[nd]).sum(0) / self.mass_np[nd].sum()
            node += list(nd)
            grp += [b] * len(nd)
            w.append(np.c_[np.ones(len(nd)), X[nd] - c])
            Mu.append(float((area[el] * np.asarray(me.mu, np.float64)[el]).sum()))
            La.append(float((area[el] * np.asarray(me.lam, np.float64)[el]).sum()))
            env.append(int(np.asarray(me.tri2env)[el[0]]))
            self.mu_np[el] = 0.0
            self.lam_np[el] = 0.0
        self.mu = self._t(self.mu_np, F64)
        self.lam = self._t(self.lam_np, F64)
        n = len(groups)
        t = self._t
        a = dict(n=n, node_np=np.asarray(node, np.int64), grp_np=np.asarray(grp, np.int64),
                 elems=np.concatenate([np.asarray(g["elems"], np.int64) for g in groups]),
                 w_np=np.concatenate(w, 0), first=np.concatenate([[0], np.cumsum([len(v) for v in w])]),
                 slaved=owner >= 0, group_elems=gel, owner=owner)
        self.affine = a
        self.an_node, self.an_grp = t(a["node_np"], I32), t(a["grp_np"], I32)
        self.an_w = t(a["w_np"], F64)
        self.ag_Mu, self.ag_La, self.ag2env = t(Mu, F64), t(La, F64), t(env, I32)
        self.ag_start = t(np.searchsorted(env, np.arange(self.B + 1)), I32)
        self.an_start = t(a["first"], I32)
        z = lambda k: torch.zeros((n, k), dtype=F64, device=self.dev)
        self.ag_q, self.ag_q0, self.ag_pq, self.ag_gq, self.ag_gq_prev, self.ag_zq = (z(6) for _ in range(6))
        self.ag_Hx, self.ag_Hy = z(9), z(9)
        self.sync_affine(None)
        self._layout = None                 # the scene layout sorts rigid triangles and sizes their LDS

    def set_affine_moduli(self, mu_full, lam_full):
        """Element moduli with rigid regions: each region's energy takes the area-weighted sums of
        its elements' moduli, and its elements stay zero in the nodal passes."""
        a = self.affine
        mu, lam = np.asarray(mu_full, np.float64).copy(), np.asarray(lam_full, np.float64).copy()
        area = np.asarray(self.me.rest_area, np.float64)
        Mu = np.array([(area[el] * mu[el]).sum() for el in a["group_elems"]])
        La = np.array([(area[el] * lam[el]).sum() for el in a["group_elems"]])
        for el in a["group_elems"]:
            mu[el] = 0.0
            lam[el] = 0.0
        self.mu_np, self.lam_np = mu, lam
        self.ag_Mu.copy_(torch.as_tensor(Mu))
        self.ag_La.copy_(torch.as_tensor(La))
        self.refresh_moduli()

    def sync_affine(self, regions=None):
        """q from the current node positions by least squares (reference implicit/affine.py:331-344).
        regions: registration indices of the regions to re-sync (None: all)."""
        a = self.affine
        if a is None or (regions is not None and len(regions) == 0):
            return
        if regions is not None:
            pos = np.empty(len(self.ag_order), np.int64)
            pos[self.ag_order] = np.arange(len(self.ag_order))
            regions = [int(pos[r]) for r in regions]
        x = self.x.double().cpu().numpy()
        q = np.zeros((a["n"], 6)) if regions is None else self.ag_q.cpu().numpy()
        for b in (range(a["n"]) if regions is None else regions):
            k = slice(a["first"][b], a["first"][b + 1])
            sol = np.linalg.lstsq(a["w_np"][k], x[a["node_np"][k]], rcond=None)[0]
            q[b, :3], q[b, 3:] = sol[:, 0], sol[:, 1]
        self.ag_q.copy_(torch.as_tensor(q))
        self.ag_pq.zero_()
        self.ag_gq.zero_()
        self.ag_gq_prev.zero_()
        v = q[a["grp_np"]]
        xs = np.stack([(a["w_np"] * v[:, 0:3]).sum(1), (a["w_np"] * v[:, 3:6]).sum(1)], 1)
        self.x[torch.as_tensor(a["node_np"], device=self.dev)] = torch.as_tensor(xs, device=self.dev)

    def affine_state(self):
        """(t, A) per rigid region in registration order."""
        qs = self.ag_q.cpu().numpy()
        q = np.empty_like(qs)
        q[self.ag_order] = qs
        return q[:, [0, 3]], np.stack([q[:, 1:3], q[:, 4:6]], 1)

    def setup_rigid_modes(self, skip_nodes=None):
        """Coarse space of every free soft piece (reference implicit/affine.py:503-537)."""
        tri = np.asarray(self.me.tri)
        a = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
        b = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
        ncomp, lab = connected_components(coo_matrix((np.ones(len(a)), (a, b)), shape=(self.N, self.N)),
                                          directed=False)
        held = np.zeros(ncomp, bool)
        held[np.unique(lab[(self.free_np == 0).any(1)])] = True
        if skip_nodes is not None and np.any(skip_nodes):
            held[np.unique(lab[np.asarray(skip_nodes, bool)])] = True
        size = np.bincount(lab, minlength=ncomp)
        keep = np.where(~held & (size >= 3))[0]
        if len(keep) == 0:
            return
        newid = np.full(ncomp, -1)
        newid[keep] = np.arange(len(keep))
        nodes = np.where(newid[lab] >= 0)[0]
        grp = newid[lab[nodes]]
        n = len(keep)
        first = np.zeros(n, np.int64)
        first[grp[::-1]] = nodes[::-1]
        # groups by environment, each group's nodes contiguous (in node order)
        genv = self.node2env_np[first]
        gorder = np.argsort(genv, kind="stable")
        rank = np.empty(n, np.int64)
        rank[gorder] = np.arange(n)
        grp = rank[grp]
        first = first[gorder]
        o = np.lexsort((nodes, grp))
        nodes, grp = nodes[o], grp[o]
        t = self._t
        self.rigid = dict(n=n)
        self.rn_node, self.rn_grp = t(nodes, I32), t(grp, I32)
        self.rg2env = t(self.node2env_np[first], I32)
        self.rg_start = t(np.searchsorted(self.node2env_np[first], np.arange(self.B + 1)), I32)
        self.rn_start = t(np.searchsorted(grp, np.arange(n + 1)), I32)
        z = lambda k: torch.zeros((n, k), dtype=F64, device=self.dev)
        self.rg_acc, self.rg_gq, self.rg_gq_prev, self.rg_zq, self.rg_H = z(3), z(3), z(3), z(3), z(9)
        self._layout = None

    # ------------------------------------------------------------------ actuation
    def set_fext(self, f):
        f = np.asarray(f, np.float64)
        if self.fext is None:
            self.fext = self._t(f, F64)
        else:
            self.fext.copy_(torch.as_tensor(f))

    def setup_pressure(self, ea, eb, ech, ch2env, cref):
        """Pressure chambers, grouped by environment (chamber order is kept within one)."""
        t = self._t
        ea, eb, ech = (np.asarray(v, np.int64) for v in (ea, eb, ech))
        ch2env = np.asarray(ch2env, np.int64)
        corder = np.argsort(ch2env, kind="stable")
        rank = np.empty(len(corder), np.int64)
        rank[corder] = np.arange(len(corder))
        ech = rank[ech]
        o = np.argsort(ech, kind="stable")
        ea, eb, ech = ea[o], eb[o], ech[o]
        self.ch_order = corder
        self.pressure = dict(n_pe=len(ea), n_ch=len(ch2env))
        self.pe_a, self.pe_b, self.pe_ch = t(ea, I32), t(eb, I32), t(ech, I32)
        self.ch2env = t(ch2env[corder], I32)
        self.ch_ref = t(np.asarray(cref, np.float64).reshape(-1, 2)[corder], F64)
        self.pe_start = t(np.searchsorted(ch2env[corder][ech], np.arange(self.B + 1)), I32)
        self.ch_p = torch.zeros(len(ch2env), dtype=F64, device=self.dev)
        self._layout = None

    def set_pressure(self, p):
        self.ch_p.copy_(torch.as_tensor(np.asarray(p, np.float64)[self.ch_order]))

    def setup_tendon(self, paths, t2env, stiff, L0):
        """Tendons (eyelet node lists), grouped by environment."""
        t = self._t
        t2env = np.asarray(t2env, np.int64)
        torder = np.argsort(t2env, kind="stable")
        self.td_order = torder
        node, prv, nxt, tid = [], [], [], []
        for i, k in enumerate(torder):
            o, n = len(node), len(paths[k])
            node += list(paths[k])
            tid += [i] * n
            prv += [-1] + list(range(o, o + n - 1))
            nxt += list(range(o + 1, o + n)) + [-1]
        self.tendon = dict(n_ey=len(node), n_td=len(t2env))
        self.ey_node, self.ey_prv, self.ey_nxt, self.ey_tid = (t(v, I32) for v in (node, prv, nxt, tid))
        self.ey_start = t(np.searchsorted(np.asarray(tid), np.arange(len(t2env) + 1)), I32)
        self.td2env = t(t2env[torder], I32)
        self.td_start = t(np.searchsorted(t2env[torder], np.arange(self.B + 1)), I32)
        self.td_stiff = t(np.asarray(stiff, np.float64)[torder], F64)
        self.td_L0 = t(np.asarray(L0, np.float64)[torder], F64)
        n = len(t2env)
        self.td_T, self.td_L, self.td_tau, self.td_keff, self.td_dL = (
            torch.zeros(n, dtype=F64, device=self.dev) for _ in range(5))

    def set_tendon(self, tension, length):
        self.td_T.copy_(torch.as_tensor(np.asarray(tension, np.float64)[self.td_order]))
        self.td_L0.copy_(torch.as_tensor(np.asarray(length, np.float64)[self.td_order]))

    def setup_springs(self, a, b, k, s2env):
        """Springs and pins (b = -1: to a world point), grouped by environment, the model's
        order kept within one (reference implicit/actuation.py:448-476)."""
        t = self._t
        s2env = np.asarray(s2env, np.int64)
        o = np.argsort(s2env, kind="stable")
        self.sp_order = o
        self.springs = dict(n=len(o))
        self.sp_a = t(np.asarray(a, np.int64)[o], I32)
        self.sp_b = t(np.asarray(b, np.int64)[o], I32)
        self.sp_k = t(np.asarray(k, np.float64)[o], F64)
        self.sp2env = t(s2env[o], I32)
        self.sp_start = t(np.searchsorted(s2env[o], np.arange(self.B + 1)), I32)
        self.sp_L0 = torch.zeros(len(o), dtype=F64, device=self.dev)
        self.sp_anchor = torch.zeros((len(o), 2), dtype=F64, device=self.dev)

    def set_springs(self, anchor, length):
        o = self.sp_order
        self.sp_anchor.copy_(torch.as_tensor(np.asarray(anchor, np.float64).reshape(-1, 2)[o]))
        self.sp_L0.copy_(torch.as_tensor(np.asarray(length, np.float64)[o]))

    def setup_activation(self, elem, ereg, n_regions):
        """Active regions: their elements' rest shape follows an activation A per region
        (reference implicit/actuation.py:340-366). The undeformed rest shape is kept on the device
        the first time."""
        if getattr(self, "Bm0", None) is None:
            self.Bm0, self.area0 = self.Bm.clone(), self.area.clone()
        self.act_elem = self._t(np.asarray(elem, np.int64), I32)
        self.act_ereg = self._t(np.asarray(ereg, np.int64), I32)
        self.activation = dict(n_el=len(elem), n_reg=int(n_regions))

    def set_activation(self, A):
        """A: (n_regions, 2, 2) activation of each region's rest shape: Bm = Bm0 A^-1 and
        area = area0 det A, in both element layouts."""
        A = np.asarray(A, np.float64).reshape(-1, 2, 2)
        Ainv = self._t(np.linalg.inv(A).reshape(-1, 4), F64)
        detA = self._t(np.linalg.det(A), F64)
        o = _lib.ElemConsts()
        o.Bm, o.area, o.mu, o.lam = (self._ptr(v) for v in (self.Bm, self.area, self.mu, self.lam))
        if self._layout is not None:
            for k in ("sc_pos", "sc_Bm", "sc_Amu", "sc_Alam", "sc_Bm32", "sc_Amu32", "sc_Alam32"):
                setattr(o, k, self._ptr(getattr(self, k)))
        L = _lib.lib()
        _lib.check(L.d2f_apply_activation_f64(self._ptr(self.act_elem), self._ptr(self.act_ereg),
                                              self.activation["n_el"], self._ptr(self.Bm0),
                                              self._ptr(self.area0), self._ptr(Ainv), self._ptr(detA),
                                              C.byref(o), self._stream()), "d2f_apply_activation_f64")
        self._act_keep = (Ainv, detA)       # alive until the launch has run

    def clear_actuators(self):
        self.pressure = None
        self.tendon = None
        self.springs = None
        self.activation = None

    # ------------------------------------------------------------------ scene-persistent layout
    def build_layout(self, max_cells=1024):
        """Per-scene triangle order (soft triangles by colour, then rigid ones), area-weighted
        element constants in float64 and float32, and the per-scene offsets the
        scene-persistent kernel reads."""
        from . import layout as LY
        me = self.me
        tri = np.asarray(me.tri, np.int64)
        T = len(tri)
        color = np.zeros(T, np.int64)
        tq = np.asarray(me.tri_quad)
        cache = {}
        for h in range(len(me.body_tri_range)):
            t0, t1 = (int(v) for v in me.body_tri_range[h])
            if t1 == t0:
                continue
            n0 = int(me.body_node_range[h][0])
            loc = tri[t0:t1] - n0
            if tq[t0] >= 0:
                color[t0:t1] = LY.color_pixel_body(tq[t0:t1], int(me.body_grid[h][1]))
            else:
                key = (loc.shape, loc.tobytes())
                if key not in cache:
                    cache[key] = LY.color_mesh(loc, int(me.body_node_range[h][1]) - n0)
                color[t0:t1] = cache[key]
        rigid = np.zeros(T, bool)
        if self.affine is not None:
            rigid[self.affine["elems"]] = True
        tenv = np.asarray(me.tri2env, np.int64)
        M = LY.MAX_COLORS + 1
        key = tenv * M + np.where(rigid, LY.MAX_COLORS, color)
        perm = np.lexsort((np.arange(T), key))
        ks = key[perm]
        t = self._t
        self.sc_perm = perm
        self.sc_tri_start = t(np.searchsorted(ks, np.arange(self.B + 1) * M), I32)
        self.sc_cs = t(np.searchsorted(ks, (np.arange(self.B)[:, None] * M + np.arange(M)[None, :]).reshape(-1)), I32)
        self.sc_tri = t(tri[perm].astype(np.int32), I32)
        Bm = self.Bm.cpu().numpy().reshape(-1, 4)[perm]       # current rest shape (activation)
        area = self.area.cpu().numpy()[perm]
        pos = np.empty(T, np.int64)
        pos[perm] = np.arange(T)
        self.sc_pos = t(pos, I32)
        Amu, Alam = area * self.mu_np[perm], area * self.lam_np[perm]
        self.sc_Bm, self.sc_Amu, self.sc_Alam = t(Bm, F64), t(Amu, F64), t(Alam, F64)
        self.sc_Bm32, self.sc_Amu32 = t(Bm.astype(np.float32), torch.float32), t(Amu.astype(np.float32), torch.float32)
        self.sc_Alam32 = t(Alam.astype(np.float32), torch.float32)
        bl = np.full(self.N, -1, np.int64)
        bn = np.asarray(me.boundary_nodes, np.int64)
        bl[bn] = np.arange(len(bn)) - np.asarray(me.boundary_starts)[np.asarray(me.boundary2env)]
        self.bloc = t(bl, I32)
        if self.ipc is not None:
            self.e_start = t(np.searchsorted(self.e2env.cpu().numpy(), np.arange(self.B + 1)), I32)
        self.max_nb = int(np.max(np.asarray(me.boundary_per_env))) if self.B else 0
        if self.pressure is not None:
            ends = np.concatenate([self.pe_a.cpu().numpy(), self.pe_b.cpu().numpy()])
            if (bl[ends] < 0).any():
                raise ValueError("a pressure edge has an end that is not a boundary node")
        self.max_cells = int(max_cells)
        self._route()
        self._plan_lds()
        self._apply_lds_overrides()
        self._layout = True

    def set_plan(self, block_size=None, max_cells=None, skin=None, lds_order=None):
        """Launch shape and LDS plan of the scene kernel (what the autotuner chooses): threads
        per workgroup (256, 512 or 1024), grid cells per scene, pair-list margin and the order
        in which the optional LDS parts are granted. Re-plans routing and LDS."""
        if block_size is not None:
            if int(block_size) not in (256, 512, 1024):
                raise ValueError("block_size must be 256, 512 or 1024")
            self.block_size = int(block_size)
        if skin is not None:
            if not float(skin) >= 0.0:
                raise ValueError("skin must be >= 0")
            self.skin = float(skin)
        if lds_order is not None:
            if sorted(lds_order) != sorted(self.LDS_ORDER):
                raise ValueError(f"lds_order must order {self.LDS_ORDER}")
            self.lds_order = tuple(lds_order)
        if max_cells is not None:
            if int(max_cells) < 1:
                raise ValueError("max_cells must be >= 1")
            self.max_cells = int(max_cells)
        if self._layout is not None:
            self._route()
            self._plan_lds()
            self._apply_lds_overrides()

    def plan(self):
        """The current launch shape and LDS plan."""
        return dict(block_size=self.block_size, max_cells=self.max_cells, skin=self.skin,
                    lds_order=list(self.lds_order), nacc_lds=self.nacc_lds, xc_lds=self.xc_lds,
                    excl_bits_lds=self.excl_bits_lds, max_pairs=self.max_pairs,
                    plain_envs=len(self.plain_envs))

    LDS_BUDGET = 150 * 1024          # bytes of LDS one scene's workgroup may take
    EXCL_LDS_MAX = 8192              # largest exclusion bitset staged in LDS, 32-bit words
    SKIN = 0.5                       # pair-list margin, in units of the contact reach
    # optional LDS parts, granted in this order while the budget lasts: per-node element
    # accumulators, positions, the bit-packed exclusion matrix, the candidate-pair list
    LDS_ORDER = ("nacc", "xc", "bits", "pairs")

    def _fixed_lds(self, nb, ag):
        """Bytes of LDS a scene workgroup cannot do without (for the current block_size), for a
        scene of nb boundary nodes and ag affine groups: block-reduction partials, boundary
        positions and accumulators, affine group state, the grid and the boundary list."""
        nw = int(self.block_size) // 64            # waves per workgroup
        red = 20 * 4 * nw + 20                     # block reductions: 20 values x 4 rows per wave, totals
        return 8 * (red + 8 * nb + 26 * ag) + 4 * (2 * self.max_cells + 1 + 2 * nb)

    def _route(self):
        """Which scenes the scene kernel takes: those whose fixed LDS fits one workgroup's budget.
        The others are stepped by the per-kernel path, in float64, in the same call (their
        indices in self.plain_envs; stats() reports them)."""
        nb = np.asarray(self.me.boundary_per_env, np.int64)
        ag = (np.diff(self.ag_start.cpu().numpy()).astype(np.int64) if self.affine is not None
              else np.zeros(self.B, np.int64))
        fits = np.array([self._fixed_lds(int(nb[e]), int(ag[e])) <= self.LDS_BUDGET for e in range(self.B)], bool)
        self.scene_envs = np.flatnonzero(fits)
        self.plain_envs = np.flatnonzero(~fits)
        if len(self.plain_envs) and not getattr(self, "_warned_route", False):
            self._warned_route = True
            warnings.warn(f"scenes {self.plain_envs[:8].tolist()}{' ...' if len(self.plain_envs) > 8 else ''} "
                          f"are too large for one workgroup's LDS: they take the per-kernel path "
                          f"(float64); the other {len(self.scene_envs)} keep the scene path.", stacklevel=4)
        self._route_t = None
        if len(self.plain_envs):
            self._route_t = dict(skip=self._t(fits.astype(np.uint8), U8),
                                 list=self._t(self._longest_first(self.scene_envs).astype(np.int32), I32))
        self._order_t = None
        if not len(self.plain_envs) and self.longest_first and self.B > self.n_cu:
            self._order_t = self._t(self._longest_first(self.scene_envs).astype(np.int32), I32)

    def _longest_first(self, envs):
        """The scenes in order of decreasing work per iteration (triangles, nodes and boundary
        nodes): workgroups are dispatched in launch order, so when there are more scenes than
        compute units the largest start first and the small ones fill in behind them."""
        me = self.me
        work = (np.diff(np.asarray(me.tri_starts, np.int64)) + 0.5 * np.diff(np.asarray(me.node_starts, np.int64))
                + np.asarray(me.boundary_per_env, np.int64))
        envs = np.asarray(envs, np.int64)
        return envs[np.argsort(-work[envs], kind="stable")]

    def _plan_lds(self):
        """Split each scene workgroup's LDS budget (for the current block_size) over the scenes
        the scene kernel takes: fixed scratch (reductions, boundary positions and accumulators,
        grid), then the optional parts in the order of lds_order while they fit: per-node
        element accumulators ("nacc"), positions ("xc"), the bit-packed exclusion matrix
        ("bits"; bit jp * n_b_e + j of a scene's bitset is its exclusion of boundary nodes jp
        and j) and the candidate-pair list ("pairs")."""
        self.excl_bits_lds, self.max_pairs, self.nacc_lds, self.xc_lds = 0, 0, 0, 0
        sub = self.scene_envs
        nn = np.diff(np.asarray(self.me.node_starts, np.int64))[sub]
        self.max_nn = int(nn.max(initial=0))
        self.sc_max_nb = int(np.asarray(self.me.boundary_per_env, np.int64)[sub].max(initial=0))
        self.max_ag = 0                            # affine groups of the largest scene: 26 doubles each
        if self.affine is not None:
            self.max_ag = int(np.diff(self.ag_start.cpu().numpy())[sub].max(initial=0))
        base = self._fixed_lds(self.sc_max_nb, self.max_ag)
        contact = not ((self.ipc is None or self.n_e == 0) and self.node is None)
        wmax = int(self.excl_words[sub].max(initial=0)) if contact else 0
        pairs_ok = False
        if contact and self.ipc is not None:   # node contact: partners come from the grid itself
            e_per = np.bincount(self.e2env.cpu().numpy(), minlength=self.B)[sub]
            pairs_ok = self.sc_max_nb < 65536 and e_per.max(initial=0) < 32768
        budget = self.LDS_BUDGET
        for part in self.lds_order:
            if part == "nacc" and base + 32 * self.max_nn <= budget:
                self.nacc_lds = 1
                base += 32 * self.max_nn
            elif part == "xc" and base + 16 * self.max_nn + 4 * 4 * self.sc_max_nb <= budget:
                self.xc_lds = 1                  # positions, if a modest pair list still fits
                base += 16 * self.max_nn
            elif part == "bits" and contact and wmax <= self.EXCL_LDS_MAX and base + 4 * wmax <= budget:
                self.excl_bits_lds = wmax
                base += 4 * wmax
            elif part == "pairs" and pairs_ok:
                self.max_pairs = int(max(0, min((budget - base) // 4, 16 * self.sc_max_nb)))
                base += 4 * self.max_pairs

    def _apply_lds_overrides(self):
        """Caps on the LDS plan, set through `lds_overrides` (a dict with any of nacc_lds,
        xc_lds, excl_bits_lds, max_pairs): each planned value is replaced by the smaller of the two.
        They select the kernel's fallbacks (colour-by-colour element pass, exclusion from the
        byte matrix, grid traversal or a pair list that overflows)."""
        for k, v in (getattr(self, "lds_overrides", None) or {}).items():
            if k not in ("nacc_lds", "xc_lds", "excl_bits_lds", "max_pairs"):
                raise KeyError(k)
            setattr(self, k, min(int(getattr(self, k)), int(v)))

    def refresh_moduli(self):
        """Re-upload element moduli after mu_np / lam_np changed (both layouts)."""
        self.mu = self._t(self.mu_np, F64)
        self.lam = self._t(self.lam_np, F64)
        if self._layout is not None:
            perm = self.sc_perm
            area = self.area.cpu().numpy()[perm]
            Amu, Alam = area * self.mu_np[perm], area * self.lam_np[perm]
            self.sc_Amu.copy_(torch.as_tensor(Amu))
            self.sc_Alam.copy_(torch.as_tensor(Alam))
            self.sc_Amu32.copy_(torch.as_tensor(Amu.astype(np.float32)))
            self.sc_Alam32.copy_(torch.as_tensor(Alam.astype(np.float32)))

    # ------------------------------------------------------------------ ABI structs
    def _ptr(self, tensor):
        return None if tensor is None else C.c_void_p(tensor.data_ptr())

    def elem_tiles_struct(self):
        """The element kernel's node tiles as the C struct (pointers into el_tiles)."""
        from .layout import tiles_struct
        return tiles_struct(self.el_tiles)

    def scene_struct(self, part=None):
        """The d2f_scene of this batch. part "plain": the per-kernel path's share of a routed
        batch (the scene kernel's scenes skipped); part "scene": the scene kernel's share."""
        s = _lib.Scene()
        s.B, s.N, s.T, s.n_b, s.n_bc = self.B, self.N, self.n_tri, self.n_b, self.n_bc
        s.n_e = self.n_e if self.ipc is not None else 0
        p = lambda name: C.c_void_p(getattr(self, name).data_ptr())
        for k in ("x", "v", "xn", "xhat", "grad", "grad_prev", "p", "diag", "diag_el", "x0", "mass",
                  "node2env", "node_start", "tri", "Bm", "area", "mu", "lam", "tri2env", "bnd", "b2e",
                  "bnd_start", "bc_idx", "bc_target", "bc_q", "gPg", "g_p", "g_Py", "y_p", "y_Py", "beta",
                  "gTp", "pHp", "alpha", "p_inf", "clr_min", "dE", "dE0", "Ener", "Eprev", "clr_ipc",
                  "q_inf", "sweep_beta", "tc", "active", "iters"):
            setattr(s, k, p(k))
        s.free_ = p("free")
        s.el = self.elem_tiles_struct()
        if self.ipc is not None:
            for k in ("ea", "eb", "ela", "elb", "e2env", "ca", "cb"):
                setattr(s, k, p(k))
            s.mu_node = self._ptr(self.mu_node)
            s.e_start, s.max_nb = p("e_start"), self.max_nb
            s.excl_bits, s.excl_bits_off = p("excl_bits"), p("excl_bits_off")
            s.excl_bits_lds = self.excl_words_max          # the per-kernel path's reading
        if self.node is not None:
            s.excl_bits, s.excl_bits_off = p("excl_bits"), p("excl_bits_off")
            s.mu_node = self._ptr(self.mu_node)
            s.max_nb = self.max_nb
        s.fext = self._ptr(self.fext)
        if self.pressure is not None:
            s.n_pe, s.n_ch = self.pressure["n_pe"], self.pressure["n_ch"]
            for k in ("pe_a", "pe_b", "pe_ch", "ch2env", "ch_ref", "ch_p"):
                setattr(s, k, p(k))
        if self.tendon is not None:
            s.n_ey, s.n_td = self.tendon["n_ey"], self.tendon["n_td"]
            for k in ("ey_node", "ey_prv", "ey_nxt", "ey_tid", "td2env", "td_stiff", "td_L0", "td_T",
                      "td_L", "td_tau", "td_keff", "td_dL"):
                setattr(s, k, p(k))
        if self.affine is not None:
            s.n_an, s.n_ag = len(self.affine["node_np"]), self.affine["n"]
            for k in ("an_node", "an_grp", "ag2env", "an_w", "ag_Mu", "ag_La", "ag_q", "ag_q0", "ag_pq",
                      "ag_gq", "ag_gq_prev", "ag_zq", "ag_Hx", "ag_Hy"):
                setattr(s, k, p(k))
        if self.rigid is not None:
            s.n_rn, s.n_rg = len(self.rn_node), self.rigid["n"]
            for k in ("rn_node", "rn_grp", "rg2env", "rg_acc", "rg_gq", "rg_gq_prev", "rg_zq", "rg_H"):
                setattr(s, k, p(k))
        if self._layout is not None:
            for k in ("sc_tri_start", "sc_cs", "sc_tri", "sc_Bm", "sc_Amu", "sc_Alam", "sc_Bm32",
                      "sc_Amu32", "sc_Alam32", "bloc", "bc_start"):
                setattr(s, k, p(k))
            if self.ipc is not None:
                s.e_start = p("e_start")
            if self.affine is not None:
                s.ag_start, s.an_start = p("ag_start"), p("an_start")
            if self.rigid is not None:
                s.rg_start, s.rn_start = p("rg_start"), p("rn_start")
            if self.pressure is not None:
                s.pe_start = p("pe_start")
            if self.tendon is not None:
                s.td_start, s.ey_start = p("td_start"), p("ey_start")
            s.max_nb, s.max_cells = (self.max_nb if part == "plain" else self.sc_max_nb), self.max_cells
            if self.excl_bits is not None:
                s.excl_bits, s.excl_bits_off = p("excl_bits"), p("excl_bits_off")
            s.excl_bits_lds, s.max_pairs = self.excl_bits_lds, self.max_pairs
            if part == "plain" or self.path == "plain":
                s.excl_bits_lds = self.excl_words_max
            s.max_nn, s.nacc_lds, s.xc_lds = self.max_nn, self.nacc_lds, self.xc_lds
            s.max_ag = self.max_ag
        if self.springs is not None:
            s.n_sp = self.springs["n"]
            for k in ("sp_a", "sp_b", "sp2env", "sp_anchor", "sp_k", "sp_L0", "sp_start"):
                setattr(s, k, p(k))
        if part is not None and self._route_t is not None:
            if part == "plain":
                s.env_skip = self._ptr(self._route_t["skip"])
            else:
                s.sc_list, s.n_list = self._ptr(self._route_t["list"]), len(self.scene_envs)
        elif (part != "plain" and self.path == "scene" and self._layout is not None
              and self._order_t is not None):
            s.sc_list, s.n_list = self._ptr(self._order_t), self.B
        s.n_col = len(self.colliders)
        for i, c in enumerate(self.colliders):
            s.col[i].kind = int(c["kind"])
            for k in ("cx", "cy", "rx", "ry", "rref", "r", "mu"):
                setattr(s.col[i], k, float(c[k]))
        return s

    def params_struct(self):
        q = _lib.Params()
        q.dt, q.h2, q.inv_dt, q.drag = self.dt, self.h2, self.inv_dt, self.drag
        q.dhat, q.kappa = float(self.cfg.ipc_d_hat), float(self.cfg.obstacle_stiffness)
        q.epsv = float(self.cfg.obstacle_friction_epsv)
        q.eps, q.tol_x, q.near_factor, q.alpha_max = self.eps, self.tol_x, self.near_factor, self.alpha_max
        q.area_shrink, q.ccd_eta, q.ls_tol, q.cg_guard = self.area_shrink, self.ccd_eta, self.ls_tol, self.cg_guard
        q.iter_max, q.conv_check = self.iter_max, self.conv_check
        q.friction, q.line_search, q.n_backtrack = int(self.friction), int(self.line_search), self.n_bt
        if self.ipc is not None:
            q.ipc_on = 1
            q.ipc_kappa, q.ipc_mu = self.ipc["kappa"], self.ipc["mu"]
            q.ipc_reach, q.ipc_gap = self.ipc["reach"], self.ipc["gap"]
            q.ipc_skin = self.skin * self.ipc["reach"]
            q.ipc_per_body = int(self.ipc["per_body"])
        if self.node is not None:
            nd = self.node
            q.node_on, q.node_fric = 1, int(nd["fric"])
            q.node_k, q.node_r, q.node_mu = nd["k"], nd["r"], nd["mu"]
            q.ipc_skin = self.skin * nd["r"]
            q.ipc_per_body = int(nd["per_body"])
        q.rigid_modes_on = int(self.rigid is not None)
        q.precision = 1 if self.precision == "mixed" else 0
        q.block_size = int(self.block_size)
        return q

    # ------------------------------------------------------------------ stepping
    def set_bc_targets(self, targets_full):
        if self.n_bc:
            t = np.asarray(targets_full, np.float64)[self.bc_idx_np]
            self.bc_target.copy_(torch.as_tensor(t))

    def step(self):
        """One implicit step of every environment (reference implicit/solver.py:433-451)."""
        L = _lib.lib()
        if self.path == "scene" and self._layout is None:
            self.build_layout()
        s, q = self.scene_struct(), self.params_struct()
        stream = C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        if self.path == "plain":
            _lib.check(L.d2f_step_plain(C.byref(s), C.byref(q), stream), "d2f_step_plain")
        elif self._route_t is None:
            if self.noalloc:
                _lib.check(L.d2f_step_scene_noalloc(C.byref(s), C.byref(q), stream), "d2f_step_scene_noalloc")
            else:
                _lib.check(L.d2f_step_scene(C.byref(s), C.byref(q), stream), "d2f_step_scene")
        else:
            if self.noalloc:
                raise ValueError("noalloc steps the scene kernel alone; this batch has scenes routed "
                                 "to the per-kernel path")
            # scenes too large for the scene kernel first, on the per-kernel path (the others'
            # nodes and scalars untouched; their affine group state is put back, since the
            # group kernels run over every group), then the scene kernel on its scenes
            keep = [getattr(self, k).clone() for k in self._GROUP_STATE if getattr(self, k, None) is not None]
            sp = self.scene_struct("plain")
            _lib.check(L.d2f_step_plain(C.byref(sp), C.byref(q), stream), "d2f_step_plain")
            if keep:
                g = self._scene_groups()
                for k, v in zip([k for k in self._GROUP_STATE if getattr(self, k, None) is not None], keep):
                    getattr(self, k)[g] = v[g]
            ss = self.scene_struct("scene")
            _lib.check(L.d2f_step_scene(C.byref(ss), C.byref(q), stream), "d2f_step_scene")

    _GROUP_STATE = ("ag_q", "ag_pq", "ag_gq", "ag_gq_prev", "ag_zq", "ag_Hx", "ag_Hy")

    def _scene_groups(self):
        """Affine groups of the scene kernel's scenes (a device index tensor)."""
        if getattr(self, "_sg", None) is None:
            env = self.ag2env.cpu().numpy()
            self._sg = torch.as_tensor(np.flatnonzero(np.isin(env, self.scene_envs)), device=self.dev)
        return self._sg

    def min_distance(self):
        """Smallest node-edge distance per environment (1e30 where none is within reach)."""
        if self.ipc is None:
            return np.full(self.B, 1.0e30)
        L = _lib.lib()
        s, q = self.scene_struct(), self.params_struct()
        stream = C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        _lib.check(L.d2f_ipc_min_distance(C.byref(s), C.byref(q), stream), "d2f_ipc_min_distance")
        return self.clr_ipc.cpu().numpy().copy()

    def ipc_pairs(self, cap=None):
        """The (edge, node) pairs the per-kernel path's contact passes evaluate at the current
        positions, sorted (d2f_ipc_pairs)."""
        if self.ipc is None:
            return np.zeros((0, 2), np.int64)
        cap = int(cap or 64 * max(self.n_e, 1))
        buf = torch.zeros(2 + 2 * cap, dtype=I32, device=self.dev)
        L = _lib.lib()
        s, q = self.scene_struct("plain"), self.params_struct()
        _lib.check(L.d2f_ipc_pairs(C.byref(s), C.byref(q), self._ptr(buf), cap, self._stream()), "d2f_ipc_pairs")
        h = buf.cpu().numpy()
        n = int(h[0])
        if n > cap:
            return self.ipc_pairs(n)
        P = h[2:2 + 2 * n].reshape(-1, 2).astype(np.int64)
        return P[np.lexsort((P[:, 1], P[:, 0]))]

    # ------------------------------------------------------------------ contact read-outs
    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def contact_force(self, friction=True, colliders=True, bodies=True):
        """Force that contact exerts ON each node, (N, 2) (reference implicit/solver.py:554-602)."""
        if getattr(self, "_fc", None) is None:
            self._fc = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
            self._fc_diag = torch.zeros((self.N, 2), dtype=F64, device=self.dev)
            self._all_free = torch.ones(self.N, dtype=U8, device=self.dev)
        L = _lib.lib()
        s, q = self.scene_struct(), self.params_struct()
        _lib.check(L.d2f_contact_force(C.byref(s), C.byref(q), int(bool(colliders)), int(bool(bodies)),
                                       int(bool(friction)), self._ptr(self._all_free), self._ptr(self._fc),
                                       self._ptr(self._fc_diag), self._stream()), "d2f_contact_force")
        return -self._fc.cpu().numpy()

    def collider_force(self, index=None, include_friction=False):
        """Reaction on the colliders per environment, (B, 2): on collider `index`, or on all of
        them together (reference implicit/solver.py:504-532, 604-612)."""
        if getattr(self, "_fcol", None) is None:
            self._fcol = torch.zeros((self.B, 2), dtype=F64, device=self.dev)
        if index is not None and not 0 <= int(index) < len(self.colliders):
            raise IndexError(f"collider {index} of {len(self.colliders)}")
        L = _lib.lib()
        s, q = self.scene_struct(), self.params_struct()
        _lib.check(L.d2f_collider_force(C.byref(s), C.byref(q), -1 if index is None else int(index),
                                        int(bool(include_friction)), self._ptr(self._fcol), self._stream()),
                   "d2f_collider_force")
        return self._fcol.cpu().numpy().copy()

    # ------------------------------------------------------------------ state I/O
    def get_x(self):
        return self.x.cpu().numpy().copy()

    def get_v(self):
        return self.v.cpu().numpy().copy()

    def set_nodes(self, start, x, v=None):
        n = len(x)
        xt = torch.as_tensor(np.asarray(x, np.float64), device=self.dev)
        self.x[start:start + n] = xt
        self.xn[start:start + n] = xt
        if v is not None:
            self.v[start:start + n] = torch.as_tensor(np.asarray(v, np.float64), device=self.dev)

    def set_v_rows(self, rows, v):
        self.v[torch.as_tensor(np.asarray(rows), device=self.dev)] = torch.as_tensor(
            np.asarray(v, np.float64), device=self.dev)

    def reset_all(self, rest):
        self.x.copy_(torch.as_tensor(np.asarray(rest, np.float64)))
        self.v.zero_()
        self.xn.zero_()
        if self.affine is not None:
            self.sync_affine(None)

    def reset_envs(self, pick):
        m = torch.as_tensor(np.asarray(pick)[self.node2env_np] != 0, device=self.dev)
        rest = torch.as_tensor(np.asarray(self.me.rest_pos, np.float64), device=self.dev)
        self.x[m] = rest[m]
        self.xn[m] = rest[m]
        self.v[m] = 0.0

    def set_collider(self, index, c):
        self.colliders[index].update(cx=c["cx"], cy=c["cy"])
