"""ctypes binding of libd2fem_hip.so (the C ABI declared in include/d2fem.h).

The library is built in-tree by diff_2d_fem_amd.build (hipcc for gfx950). There is no CPU
fallback: if the library is missing, every solve raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_NAME = "libd2fem_hip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

MAX_COLLIDERS = 8
NOALLOC_MAX_NODES = 131072     # nodes of d2f_step_scene_noalloc's static work arrays
MAX_COLORS = 24
P = C.c_void_p


class ElemTiles(C.Structure):
    _fields_ = [("node_off", C.c_void_p), ("nodes", C.c_void_p), ("halo_off", C.c_void_p),
                ("tri_local", C.c_void_p), ("hnode", C.c_void_p), ("hnode_off", C.c_void_p),
                ("hdest", C.c_void_p), ("part", C.c_void_p), ("max_nodes", C.c_int32),
                ("n_hnodes", C.c_int32)]


class Collider(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32)] + [
        (k, C.c_double) for k in ("cx", "cy", "rx", "ry", "rref", "r", "mu")]


_SCENE_FIELDS = (
    [(k, C.c_int32) for k in ("B", "N", "T", "n_b", "n_e", "n_bc")]
    + [(k, P) for k in ("x", "v", "xn", "xhat", "grad", "grad_prev", "p", "diag", "diag_el", "x0",
                        "mass", "free_", "node2env", "node_start",
                        "tri", "Bm", "area", "mu", "lam", "tri2env",
                        "bnd", "b2e", "bnd_start", "ea", "eb", "ela", "elb", "e2env", "ca", "cb",
                        "mu_node",
                        "bc_idx", "bc_target", "bc_q",
                        "gPg", "g_p", "g_Py", "y_p", "y_Py", "beta", "gTp", "pHp", "alpha", "p_inf",
                        "clr_min", "dE", "dE0", "Ener", "Eprev", "clr_ipc", "q_inf", "sweep_beta", "tc",
                        "active", "iters", "fext")]
    + [("n_pe", C.c_int32), ("n_ch", C.c_int32)]
    + [(k, P) for k in ("pe_a", "pe_b", "pe_ch", "ch2env", "ch_ref", "ch_p")]
    + [("n_ey", C.c_int32), ("n_td", C.c_int32)]
    + [(k, P) for k in ("ey_node", "ey_prv", "ey_nxt", "ey_tid", "td2env", "td_stiff", "td_L0",
                        "td_T", "td_L", "td_tau", "td_keff", "td_dL")]
    + [("n_an", C.c_int32), ("n_ag", C.c_int32)]
    + [(k, P) for k in ("an_node", "an_grp", "ag2env", "an_w", "ag_Mu", "ag_La", "ag_q", "ag_q0",
                        "ag_pq", "ag_gq", "ag_gq_prev", "ag_zq", "ag_Hx", "ag_Hy")]
    + [("n_rn", C.c_int32), ("n_rg", C.c_int32)]
    + [(k, P) for k in ("rn_node", "rn_grp", "rg2env", "rg_acc", "rg_gq", "rg_gq_prev", "rg_zq", "rg_H")]
    + [("n_col", C.c_int32), ("pad_", C.c_int32), ("col", Collider * MAX_COLLIDERS)]
    + [(k, P) for k in ("sc_tri_start", "sc_cs", "sc_tri", "sc_Bm", "sc_Amu", "sc_Alam", "sc_Bm32",
                        "sc_Amu32", "sc_Alam32", "bloc", "e_start", "bc_start", "ag_start", "an_start",
                        "rg_start", "rn_start", "pe_start", "td_start", "ey_start")]
    + [("max_nb", C.c_int32), ("max_cells", C.c_int32)]
    + [("excl_bits", P), ("excl_bits_off", P), ("excl_bits_lds", C.c_int32), ("max_pairs", C.c_int32)]
    + [("max_nn", C.c_int32), ("nacc_lds", C.c_int32), ("xc_lds", C.c_int32), ("pad3_", C.c_int32)]
    + [("el", ElemTiles)]
    + [("n_sp", C.c_int32), ("max_ag", C.c_int32)]
    + [(k, P) for k in ("sp_a", "sp_b", "sp2env", "sp_anchor", "sp_k", "sp_L0", "sp_start")]
    + [("env_skip", P), ("sc_list", P), ("n_list", C.c_int32), ("pad4_", C.c_int32)]
)


class ElemConsts(C.Structure):
    _fields_ = [(k, P) for k in ("Bm", "area", "mu", "lam", "sc_pos", "sc_Bm", "sc_Amu", "sc_Alam",
                                 "sc_Bm32", "sc_Amu32", "sc_Alam32")]


class Minres(C.Structure):
    _fields_ = ([("B", C.c_int32), ("maxiter", C.c_int32)]
                + [(k, C.c_void_p) for k in ("node_start", "mass", "fmask", "dinv", "b", "x", "v", "r1", "r2",
                                             "w1", "w2", "Kv", "scal", "status")]
                + [("rtol", C.c_double)])


class Scene(C.Structure):
    _fields_ = _SCENE_FIELDS


class Params(C.Structure):
    _fields_ = ([(k, C.c_double) for k in ("dt", "h2", "inv_dt", "drag", "dhat", "kappa", "epsv",
                                           "eps", "tol_x", "near_factor", "alpha_max", "area_shrink",
                                           "ccd_eta", "ls_tol", "cg_guard", "ipc_kappa", "ipc_mu",
                                           "ipc_reach", "ipc_gap", "ipc_skin")]
                + [(k, C.c_int32) for k in ("iter_max", "conv_check", "friction", "line_search",
                                            "n_backtrack", "ipc_on", "ipc_per_body", "rigid_modes_on",
                                            "precision", "block_size", "node_on", "node_fric")]
                + [(k, C.c_double) for k in ("node_k", "node_r", "node_mu")])


_lib = None


class NativeError(RuntimeError):
    """A call into the HIP library failed."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("D2FEM_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m diff_2d_fem_amd.build` "
                           f"(hipcc, gfx950). There is no CPU fallback.")
    L = C.CDLL(str(path))
    L.d2f_abi_version.restype = C.c_int
    L.d2f_last_error.restype = C.c_char_p
    L.d2f_struct_sizes.argtypes = [C.POINTER(C.c_int64)]
    L.d2f_step_plain.argtypes = [C.POINTER(Scene), C.POINTER(Params), P]
    L.d2f_ipc_min_distance.argtypes = [C.POINTER(Scene), C.POINTER(Params), P]
    L.d2f_ipc_pairs.argtypes = [C.POINTER(Scene), C.POINTER(Params), P, C.c_int32, P]
    if hasattr(L, "d2f_step_scene"):
        L.d2f_step_scene.argtypes = [C.POINTER(Scene), C.POINTER(Params), P]
        L.d2f_step_scene_noalloc.argtypes = [C.POINTER(Scene), C.POINTER(Params), P]
    i32, f64 = C.c_int32, C.c_double
    L.d2f_elastic_grad_diag_f64.argtypes = [P, P, P, P, P, P, i32, f64, i32, C.POINTER(ElemTiles), P, P, P]
    L.d2f_elastic_pHp_f64.argtypes = [P, P, P, P, P, P, P, P, i32, f64, i32, P, P]
    L.d2f_inversion_alpha_f64.argtypes = [P, P, P, P, i32, f64, P, P]
    L.d2f_elastic_energy_f64.argtypes = [P, P, P, P, P, P, P, i32, f64, P, P]
    L.d2f_elastic_hvp_f64.argtypes = [P, P, P, P, P, P, P, i32, f64, P, P, P]
    L.d2f_elastic_pressure_hvp_f64.argtypes = [P, P, P, P, P, P, i32, f64, P, P, P, P, P, P, P]
    L.d2f_collider_force.argtypes = [C.POINTER(Scene), C.POINTER(Params), i32, i32, P, P]
    L.d2f_contact_force.argtypes = [C.POINTER(Scene), C.POINTER(Params), i32, i32, i32, P, P, P, P]
    L.d2f_collider_hvp_f64.argtypes = [P, P, P, i32, P, i32, i32, f64, f64, f64, P, P]
    L.d2f_collider_grad_diag_f64.argtypes = [P, P, i32, P, i32, i32, f64, f64, f64, i32, P, P, P]
    i64 = C.c_int64
    for name in ("d2f_unique_pairs", "d2f_unique_triples"):
        getattr(L, name).restype = i64
    L.d2f_unique_pairs.argtypes = [P, i64, P, P, P, P]
    L.d2f_unique_triples.argtypes = [P, i64, P, P, P]
    L.d2f_exclusion_csr.restype = i64
    L.d2f_exclusion_csr.argtypes = [P, i64, i64, P, P, i64, i32, f64, P, P, i64]
    L.d2f_apply_activation_f64.argtypes = [P, P, i32, P, P, P, P, C.POINTER(ElemConsts), P]
    L.d2f_node_grad_diag_f64.argtypes = [P, P, P, i32, P, P, f64, f64, f64, i32, P, P, P]
    L.d2f_node_hvp_f64.argtypes = [P, P, P, P, i32, P, P, f64, f64, f64, P, P]
    L.d2f_elastic_drho_f64.argtypes = [P, P, P, P, P, P, P, i32, f64, f64, P, P]
    L.d2f_pressure_hvp_f64.argtypes = [P, P, P, P, P, i32, f64, P, P]
    L.d2f_pressure_dp_f64.argtypes = [P, P, P, P, P, i32, f64, P, P]
    L.d2f_ipc_hvp_f64.argtypes = [P, P, P, i32, P, P, P, P, f64, f64, f64, P, P]
    L.d2f_ipc_grad_diag_f64.argtypes = [P, P, i32, P, P, P, P, f64, f64, f64, P, P, P]
    L.d2f_affine_expand_f64.argtypes = [P, P, P, P, P, i32, P, P]
    L.d2f_affine_reduce_f64.argtypes = [P, P, P, P, P, P, P, P, i32, i32, i32, P, P]
    L.d2f_minres_begin_f64.argtypes = [C.POINTER(Minres), P]
    L.d2f_minres_step_f64.argtypes = [C.POINTER(Minres), P]
    sizes = (C.c_int64 * 4)()
    L.d2f_struct_sizes(sizes)
    want = (C.sizeof(Scene), C.sizeof(Params), C.sizeof(Collider), C.sizeof(Minres))
    if tuple(sizes) != want:
        raise RuntimeError(f"{path}: struct sizes {tuple(sizes)} do not match the binding {want}; "
                           f"rebuild the library")
    if L.d2f_abi_version() != 2:
        raise RuntimeError(f"{path}: ABI version {L.d2f_abi_version()}, expected 2")
    _lib = L
    return L


D2F_ECAP = -3          # a scene exceeds a compiled capacity of the scene-persistent path


def check(rc, what):
    if rc != 0:
        msg = lib().d2f_last_error().decode(errors="replace")
        err = NativeError(f"{what} failed ({rc}): {msg}")
        err.code = int(rc)
        raise err
