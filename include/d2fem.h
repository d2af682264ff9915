/* d2fem.h — C ABI of the MI355X implicit PNCG-IPC step (libd2fem_hip.so).
 *
 * The library replaces the device side of Diff_2D_FEM's ImplicitIntegrator.simulate(model, dt):
 * the Warp kernel sequence PNCG2DSolver.step_targets / _iterate (implicit/solver.py:252-451)
 * and the kernels it launches (implicit/kernels.py, core/sdf.py, implicit/ipc_contact.py,
 * implicit/affine.py, implicit/actuation.py). The reference has no native FFI for this path —
 * its Python calls Warp kernels directly — so each entry point below names the reference
 * function or kernel family it stands in for.
 *
 * Conventions
 *   - Every pointer in d2f_scene is DEVICE memory owned by the caller (the Python layer keeps
 *     it in PyTorch-ROCm tensors). The library allocates nothing and keeps no state between
 *     calls.
 *   - 2-vectors are interleaved (x, y) float64 pairs; 2x2 matrices are row-major 4-tuples.
 *   - Calls are asynchronous on `stream` (a hipStream_t passed as void*), except where the
 *     comment says the call reads device results back (the per-kernel path's convergence poll).
 *   - Return 0 on success, a negative D2F_E* code on failure; d2f_last_error() describes the
 *     most recent failure of the calling thread. Nothing throws or aborts across the ABI.
 */
#ifndef D2FEM_H
#define D2FEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D2F_ABI_VERSION 2

#define D2F_OK 0
#define D2F_EINVAL (-1)   /* bad argument or inconsistent sizes */
#define D2F_EHIP (-2)     /* a HIP runtime call failed */
#define D2F_ECAP (-3)     /* a scene exceeds a compiled capacity of the scene-persistent path */

#define D2F_MAX_COLLIDERS 8
#define D2F_MAX_COLORS 24
#define D2F_ELEM_CHUNK 1024       /* triangles per workgroup of the tiled element kernel */
#define D2F_ELEM_NODES_MAX 1024   /* longest node list one chunk sums in LDS (10-bit local indices) */

/* Node tiles of the element kernel. Triangle t belongs to chunk t / D2F_ELEM_CHUNK; chunk c
 * touches the nodes nodes[node_off[c] .. node_off[c+1]): first the nodes it owns (every
 * triangle at them is in chunk c), then its halo nodes (shared with other chunks), each part
 * ascending. tri_local[t] holds the positions of triangle t's nodes in its chunk's list, 10 bits
 * each (bits 0-9, 10-19, 20-29). The chunk adds into its owned nodes directly and writes its sums
 * at its halo nodes, halo entries s = halo_off[c] .. halo_off[c+1], into part[hdest[s]] (4
 * doubles each: grad x, grad y, diag x, diag y), where part is ordered by node; a second pass
 * adds, for each halo node hnode[k], part[hnode_off[k] .. hnode_off[k+1]) in chunk order: no
 * global atomics (within a chunk the sums are LDS atomics).
 * part is workspace: launches that share one tiles struct must not overlap. max_nodes = the
 * longest list; a list longer than D2F_ELEM_NODES_MAX makes the launch take the untiled kernel. */
typedef struct d2f_elem_tiles {
  const int32_t *node_off;            /* n_chunks + 1 */
  const int32_t *nodes;               /* node_off[n_chunks] */
  const int32_t *halo_off;            /* n_chunks + 1 */
  const uint32_t *tri_local;          /* T */
  const int32_t *hnode;               /* n_hnodes */
  const int32_t *hnode_off;           /* n_hnodes + 1 */
  const int32_t *hdest;               /* halo_off[n_chunks] */
  double *part;                       /* 4 halo_off[n_chunks] */
  int32_t max_nodes, n_hnodes;
} d2f_elem_tiles;

/* A rigid collider (reference sim_model.py:167-174; kinds core/objects.py:22-24):
 * 0 disc, 1 ellipse, 2 rounded box, 3 rounded triangle, 4 half-space (rx, ry = unit normal). */
typedef struct d2f_collider {
  int32_t kind, pad_;
  double cx, cy, rx, ry, rref, r, mu;
} d2f_collider;

/* The scene: mesh, state and term data of every environment, laid out flat with per-scene
 * offsets (CSR). Sizes are element counts. */
typedef struct d2f_scene {
  int32_t B, N, T, n_b, n_e, n_bc;
  /* nodes [N] */
  double *x, *v, *xn, *xhat, *grad, *grad_prev, *p, *diag, *diag_el, *x0; /* 2N each */
  /* grad and grad_prev are working arrays: d2f_step_scene swaps their roles every iteration,
   * so after it either may hold the last gradient */
  double *mass;                       /* N */
  const uint8_t *free_;               /* N: 1 free, 0 kinematically driven */
  const int32_t *node2env;            /* N */
  const int32_t *node_start;          /* B+1 */
  /* triangles [T] (reference tri order) */
  const int32_t *tri;                 /* 3T */
  const double *Bm;                   /* 4T, D_m^-1 row-major */
  const double *area, *mu, *lam;      /* T (mu, lam zeroed on rigid regions) */
  const int32_t *tri2env;             /* T */
  /* boundary nodes [n_b] and edges [n_e] */
  const int32_t *bnd, *b2e;           /* n_b */
  const int32_t *bnd_start;           /* B+1 */
  const int32_t *ea, *eb, *ela, *elb, *e2env; /* n_e */
  const double *ca, *cb;              /* n_e: vertex-correction shares */
  const double *mu_node;              /* N per-body friction, or NULL */
  /* kinematic nodes [n_bc] */
  const int32_t *bc_idx;              /* n_bc */
  double *bc_target;                  /* 2 n_bc */
  double *bc_q;                       /* 2N sweep direction (zero off the kinematic nodes) */
  /* per-environment scalars [B] (float64 accumulators) */
  double *gPg, *g_p, *g_Py, *y_p, *y_Py, *beta, *gTp, *pHp, *alpha, *p_inf, *clr_min;
  double *dE, *dE0, *Ener, *Eprev, *clr_ipc, *q_inf, *sweep_beta, *tc;
  int32_t *active, *iters;
  /* external force [N] (gravity + applied), or NULL */
  const double *fext;                 /* 2N */
  /* pressure edges [n_pe] */
  int32_t n_pe, n_ch;
  const int32_t *pe_a, *pe_b, *pe_ch, *ch2env; /* n_pe, n_pe, n_pe, n_ch */
  const double *ch_ref;               /* 2 n_ch */
  const double *ch_p;                 /* n_ch */
  /* tendons: eyelets [n_ey], tendons [n_td] */
  int32_t n_ey, n_td;
  const int32_t *ey_node, *ey_prv, *ey_nxt, *ey_tid, *td2env;
  const double *td_stiff, *td_L0, *td_T;
  double *td_L, *td_tau, *td_keff, *td_dL;
  /* affine bodies: slaved nodes [n_an], groups [n_ag] */
  int32_t n_an, n_ag;
  const int32_t *an_node, *an_grp, *ag2env;
  const double *an_w;                 /* 3 n_an: (1, X - c) */
  const double *ag_Mu, *ag_La;        /* n_ag */
  double *ag_q, *ag_q0, *ag_pq, *ag_gq, *ag_gq_prev, *ag_zq; /* 6 n_ag */
  double *ag_Hx, *ag_Hy;              /* 9 n_ag */
  /* rigid-mode groups: member nodes [n_rn], groups [n_rg] */
  int32_t n_rn, n_rg;
  const int32_t *rn_node, *rn_grp, *rg2env;
  double *rg_acc, *rg_gq, *rg_gq_prev, *rg_zq; /* 3 n_rg */
  double *rg_H;                       /* 9 n_rg */
  /* colliders */
  int32_t n_col, pad_;
  d2f_collider col[D2F_MAX_COLLIDERS];
  /* ---- scene-persistent path: per-scene layout (all grouped by environment) ---- */
  /* triangles of scene e, soft ones first by colour (no two triangles of one colour share a
   * node), then the rigid-region ones: colour c of scene e is [sc_cs[e*(D2F_MAX_COLORS+1)+c],
   * sc_cs[e*(D2F_MAX_COLORS+1)+c+1]), the rigid tail starts at entry D2F_MAX_COLORS and ends at
   * sc_tri_start[e+1]. Element constants are premultiplied by the rest area. */
  const int32_t *sc_tri_start;        /* B+1 */
  const int32_t *sc_cs;               /* B*(D2F_MAX_COLORS+1) */
  const int32_t *sc_tri;              /* 3T */
  const double *sc_Bm, *sc_Amu, *sc_Alam;   /* 4T, T, T (float64 path) */
  const float *sc_Bm32, *sc_Amu32, *sc_Alam32; /* 4T, T, T (mixed path) */
  const int32_t *bloc;                /* N: index among its scene's boundary nodes, or -1 */
  const int32_t *e_start;             /* B+1 boundary edges per scene */
  const int32_t *bc_start;            /* B+1 (bc_idx grouped by environment) */
  const int32_t *ag_start, *an_start; /* B+1 groups per scene, n_ag+1 slaved nodes per group */
  const int32_t *rg_start, *rn_start; /* B+1, n_rg+1 */
  const int32_t *pe_start;            /* B+1 pressure edges per scene */
  const int32_t *td_start, *ey_start; /* B+1 tendons per scene, n_td+1 eyelets per tendon */
  int32_t max_nb, max_cells;          /* largest per-scene boundary count; grid cell cap */
  /* per-scene self-contact exclusion, bit-packed (bit jp*n_b_e + j of scene e's words: boundary
   * nodes jp and j do not contact), staged in LDS when excl_bits_lds is set; pair-list
   * capacity per scene (0: traverse the grid each pass) */
  const uint32_t *excl_bits;          /* sum_e ceil(n_b_e^2 / 32) words */
  const int64_t *excl_bits_off;       /* B+1 word offsets */
  int32_t excl_bits_lds, max_pairs;   /* words of LDS for the bitset (0: read it from global); pairs cap.
                                       * d2f_step_plain reads excl_bits_lds as the largest scene's
                                       * bitset in words (staged in LDS by its contact pass when
                                       * it fits there) */
  /* largest per-scene node count; nacc_lds = 1: the element pass accumulates every node's
   * gradient and diagonal in LDS (32 bytes per node), 0: colour-by-colour read-modify-write */
  int32_t max_nn, nacc_lds;
  /* xc_lds = 1: the scene's positions are also kept in LDS (16 bytes per node), written
   * through on every update and read from there */
  int32_t xc_lds, pad3_;
  /* per-kernel path: node tiles of the element kernel (node_off NULL: atomics only) */
  d2f_elem_tiles el;
  /* springs and pins [n_sp], grouped by environment (reference implicit/actuation.py:368-476):
   * E = k/2 (|x_a - x_b| - L0)^2 with x_b = sp_anchor when sp_b < 0 (a spring to the world);
   * max_ag: largest per-scene affine group count, required (> 0) when n_ag > 0 (the scene
   * kernel keeps their coordinates, direction and moduli in LDS, 26 doubles per group) */
  int32_t n_sp, max_ag;
  const int32_t *sp_a, *sp_b, *sp2env;  /* n_sp */
  const double *sp_anchor;              /* 2 n_sp */
  const double *sp_k, *sp_L0;           /* n_sp */
  const int32_t *sp_start;              /* B+1 springs per scene (scene path) */
  /* routing of a batch between the two paths (both NULL: every scene on the called path).
   * env_skip[e] = 1: d2f_step_plain leaves scene e untouched (its nodes, targets, scalars and
   * iteration count); sc_list: d2f_step_scene steps only the n_list scenes listed, one
   * workgroup each (the LDS plan fields then describe the listed scenes). */
  const uint8_t *env_skip;              /* B, or NULL */
  const int32_t *sc_list;               /* n_list, or NULL */
  int32_t n_list, pad4_;
} d2f_scene;

/* Solver options and step constants (reference PNCG2DSolver.__init__, implicit/solver.py:97-226,
 * and ImplicitIntegrator._build, implicit/integrator.py:252-311). */
typedef struct d2f_params {
  double dt, h2, inv_dt, drag;
  double dhat, kappa, epsv;                 /* collider barrier and friction smoothing */
  double eps, tol_x, near_factor, alpha_max, area_shrink, ccd_eta;
  double ls_tol, cg_guard;
  double ipc_kappa, ipc_mu, ipc_reach, ipc_gap;
  double ipc_skin;                          /* scene path: pair-list margin beyond the reach */
  int32_t iter_max, conv_check, friction, line_search, n_backtrack;
  int32_t ipc_on, ipc_per_body, rigid_modes_on, precision; /* precision: 0 float64, 1 mixed */
  int32_t block_size;                        /* scene path: threads per workgroup (256/512/1024) */
  /* "node" contact between bodies (reference implicit/kernels.py:755-869,
   * implicit/body_contact.py): penalty node_k (node_r - d) between boundary nodes of one scene
   * closer than node_r, not excluded against each other; node_fric adds Coulomb friction on the
   * same pairs with coefficient node_mu (per body through scene mu_node when ipc_per_body). Uses
   * the scene's excl_bits / excl_bits_off; exclusive with ipc_on. ipc_skin is the scene path's grid
   * margin for either contact. */
  int32_t node_on, node_fric;
  double node_k, node_r, node_mu;
} d2f_params;

/* ---- library ---------------------------------------------------------------------------- */
int d2f_abi_version(void);
/* Text of the calling thread's last failure (empty if none). */
const char *d2f_last_error(void);
/* sizeof(d2f_scene), sizeof(d2f_params), sizeof(d2f_collider), sizeof(d2f_minres) into out[0..3],
 * so a binding can check its mirror of the structs. */
void d2f_struct_sizes(int64_t *out);

/* ---- per-kernel entry points (float64), one per reference kernel family ------------------ */
/* grad += h2 dPsi/dx, diag += h2 max(diag d2Psi/dx2, 0) over all triangles.
 * Replaces grad_diag_elastic_kernel (implicit/kernels.py:86-128). With tiles (see
 * d2f_elem_tiles; max_nodes <= D2F_ELEM_NODES_MAX) chunk c gathers its nodes' positions into
 * LDS, sums its triangles' contributions there and adds them once per node it owns (plain
 * read-modify-write). Its sums at halo nodes are atomics below 4,194,304 triangles; from there
 * on they go to the tiles' partial slots, which a second launch adds per node in chunk order
 * (no global atomics). tiles NULL: one atomic per contribution. */
int d2f_elastic_grad_diag_f64(const double *x, const int32_t *tri, const double *Bm,
                              const double *area, const double *mu, const double *lam, int32_t T,
                              double h2, int32_t clamp, const d2f_elem_tiles *tiles,
                              double *grad, double *diag, void *stream);
/* pHp[env] += h2 max(p^T d2Psi p, 0) per triangle. Replaces pHp_elastic_kernel
 * (implicit/kernels.py:259-273). */
int d2f_elastic_pHp_f64(const double *x, const double *p, const int32_t *tri, const double *Bm,
                        const double *area, const double *mu, const double *lam,
                        const int32_t *tri2env, int32_t T, double h2, int32_t clamp, double *pHp,
                        void *stream);
/* alpha[env] = min(alpha, smallest positive area-shrink root). Replaces inversion_alpha_kernel
 * (implicit/kernels.py:304-341). */
int d2f_inversion_alpha_f64(const double *x, const double *p, const int32_t *tri,
                            const int32_t *tri2env, int32_t T, double shrink, double *alpha,
                            void *stream);
/* E[env] += h2 A Psi per triangle. Replaces energy_elastic_kernel (implicit/kernels.py:454-468). */
int d2f_elastic_energy_f64(const double *x, const int32_t *tri, const double *Bm,
                           const double *area, const double *mu, const double *lam,
                           const int32_t *tri2env, int32_t T, double h2, double *E, void *stream);
/* y += h2 (d2Psi/dx2) v, unclamped. Replaces hvp_elastic_kernel (implicit/adjoint.py:70-85).
 * tiles as for d2f_elastic_grad_diag_f64 (chunk c sums its triangles' products in LDS and adds
 * each node it owns once; halo sums by atomics, or from 4,194,304 triangles on through the
 * tiles' partial slots and a second launch); tiles NULL: one atomic per contribution. Calls
 * that share one tiles struct must not overlap (its partial slots are scratch). */
int d2f_elastic_hvp_f64(const double *x, const double *v, const int32_t *tri, const double *Bm,
                        const double *area, const double *mu, const double *lam, int32_t T,
                        double h2, const d2f_elem_tiles *tiles, double *y, void *stream);
/* The same product plus the pressure chambers' (d2f_pressure_hvp_f64's term) in one tiled pass:
 * chunk c of the tiles also takes the chamber edges pe_off[c] .. pe_off[c + 1], each with its two
 * nodes as entries (lo | hi << 10) of the chunk's node list and its chamber's index into ch_p.
 * No counterpart in the reference (its adjoint refuses pressure chambers); tiles required. */
int d2f_elastic_pressure_hvp_f64(const double *x, const double *v, const double *Bm,
                                 const double *area, const double *mu, const double *lam,
                                 int32_t T, double h2, const d2f_elem_tiles *tiles,
                                 const int32_t *pe_off, const uint32_t *pe_loc,
                                 const int32_t *pe_ch, const double *ch_p, double *y, void *stream);

/* ---- whole implicit steps ----------------------------------------------------------------- */
/* One implicit step of every environment on the per-kernel path: one launch per reference
 * kernel family, float64, atomics for scatters; polls convergence every conv_check iterations
 * (a device->host read). Replaces PNCG2DSolver.step_targets (implicit/solver.py:433-451). */
int d2f_step_plain(const d2f_scene *s, const d2f_params *prm, void *stream);
/* Smallest node-edge distance per environment into s->clr_ipc (1e30 where no pair is within
 * reach); uses s->x0 as scratch. Replaces IPCContact.min_distance (implicit/ipc_contact.py:441-447),
 * the "bodies must start apart" check of ImplicitIntegrator._build. */
int d2f_ipc_min_distance(const d2f_scene *s, const d2f_params *prm, void *stream);

/* The node-edge pairs the per-kernel path's contact passes evaluate at s->x: every (boundary
 * edge k, boundary node n) of one scene with n not an end of k, not excluded against either
 * end, and closer than prm->ipc_reach to the edge (the candidates of IPCContact's hash-grid
 * query after its filters, implicit/ipc_contact.py:212-235). pairs[0] receives the count;
 * pairs[2 + 2i], pairs[3 + 2i] the edge and node of pair i < cap, in no particular order. */
int d2f_ipc_pairs(const d2f_scene *s, const d2f_params *prm, int32_t *pairs, int32_t cap, void *stream);

/* The same step on the scene-persistent path: one workgroup per environment runs the whole
 * iteration loop in one launch, with no host synchronisation; prm->precision 0 = float64,
 * 1 = mixed (float32 element arithmetic, float64 state and reductions); prm->block_size picks
 * the workgroup size. Needs the sc_* layout fields. Returns D2F_ECAP when a scene's boundary
 * does not fit the LDS budget (use d2f_step_plain for it). */
int d2f_step_scene(const d2f_scene *s, const d2f_params *prm, void *stream);

/* ---- read-outs ------------------------------------------------------------------------------ */
/* Reaction on the colliders per environment into F[2B]: collider `index`, or all (index < 0);
 * the barrier, plus friction when `friction` (and prm->friction) is set. Replaces
 * PNCG2DSolver.collider_force (implicit/solver.py:504-532, 604-612). */
int d2f_collider_force(const d2f_scene *s, const d2f_params *prm, int32_t index, int32_t friction,
                       double *F, void *stream);
/* MINUS the contact force on every node into g[2N] (dg[2N] receives the passes' diagonal):
 * colliders' barrier (+friction), IPC contact between bodies (+friction), with h^2 = 1 and
 * all_free[N] (all ones) in place of the free mask, so kinematic nodes are included. Replaces
 * PNCG2DSolver.contact_force (implicit/solver.py:554-602). */
int d2f_contact_force(const d2f_scene *s, const d2f_params *prm, int32_t colliders, int32_t bodies,
                      int32_t friction, const uint8_t *all_free, double *g, double *dg, void *stream);

/* ---- scene assembly (host memory, host code) -------------------------------------------- */
/* The distinct unordered pairs among the rows of E[m][2]: each row sorted, the distinct rows in
 * lexicographic order into key[u][2], with the row each first occurs at (first[u]), how many rows
 * equal it (count[u]) and, when inverse is not NULL, each row's pair (inverse[row]). Returns the
 * number of distinct pairs (-1: bad argument). Replaces the edge tables of core/mesh.py:126-131
 * (_finish's boundary edges), 177 (_material_excl) and 247-248 (check_mesh). */
int64_t d2f_unique_pairs(const int64_t *E, int64_t m, int64_t *key, int64_t *first, int64_t *count,
                         int64_t *inverse);
/* The same for rows of three (the duplicate-triangle check of core/mesh.py:224). */
int64_t d2f_unique_triples(const int64_t *R, int64_t m, int64_t *key, int64_t *first, int64_t *count);
/* Self-contact exclusion of one body as sorted lists over its boundary nodes (CSR): boundary
 * node s (of boundary[0..nb)) excludes every boundary node within k mesh-edge hops, itself
 * included (the breadth-first search of core/mesh.py:133-153), and with radius > 0 every one
 * closer than radius along the mesh's edges at rest (pos: rest positions; the shortest paths
 * of _material_excl, core/mesh.py:163-186, stopped at the radius); the relation is made
 * symmetric (core/mesh.py:157-158). The same sets as the reference's nb x nb matrix, which is
 * never formed. row_ptr[nb + 1] and cols (capacity cap) receive the lists; returns the number of
 * entries (more than cap: only cap written, call again with a larger cols), or < 0 on bad input. */
int64_t d2f_exclusion_csr(const int64_t *tri, int64_t T, int64_t n_nodes, const double *pos,
                          const int64_t *boundary, int64_t nb, int32_t k, double radius, int64_t *row_ptr,
                          int32_t *cols, int64_t cap);

/* ---- active strain ------------------------------------------------------------------------ */
/* Element constants an activation rewrites: the reference-order arrays and, when sc_pos is not
 * NULL, the scene layout (element t sits at sc_pos[t]; constants premultiplied by the area, in
 * float64 and float32). */
typedef struct d2f_elem_consts {
  double *Bm, *area;                   /* 4T, T */
  const double *mu, *lam;              /* T, current moduli */
  const int32_t *sc_pos;               /* T, or NULL */
  double *sc_Bm, *sc_Amu, *sc_Alam;    /* 4T, T, T */
  float *sc_Bm32, *sc_Amu32, *sc_Alam32;
} d2f_elem_consts;
/* Bm[t] = Bm0[t] Ainv[r], area[t] = area0[t] detA[r] for each element t = elem[j] of region
 * r = ereg[j] (and the scene layout's copies). Replaces apply_activation_kernel
 * (implicit/actuation.py:328-337). */
int d2f_apply_activation_f64(const int32_t *elem, const int32_t *ereg, int32_t n_el, const double *Bm0,
                             const double *area0, const double *Ainv, const double *detA,
                             const d2f_elem_consts *out, void *stream);

/* ---- adjoint (float64) ----------------------------------------------------------------------- */
/* y += h2 H_contact v at x for the selected colliders (which: 0 discs, 1 the other shapes through
 * their level sets, 2 all), barrier gate frozen at x, unclamped. Replaces chvp_disc_kernel and
 * chvp_sdf_kernel (implicit/adjoint.py:89-112). */
int d2f_collider_hvp_f64(const double *x, const double *v, const int32_t *bnd, int32_t n_b,
                         const d2f_collider *cols, int32_t n_col, int32_t which, double dhat,
                         double kappa, double h2, double *y, void *stream);
/* grad += h2 kappa b' grad phi, diag += h2 curvature of the selected colliders (the contact
 * part of the adjoint's Jacobi preconditioner). Replaces grad_diag_disc_kernel
 * (implicit/kernels.py:132-157) and grad_diag_sdf_kernel (core/sdf.py:141-154) as the adjoint
 * calls them (implicit/adjoint.py:356-370). */
int d2f_collider_grad_diag_f64(const double *x, const int32_t *bnd, int32_t n_b,
                               const d2f_collider *cols, int32_t n_col, int32_t which, double dhat,
                               double kappa, double h2, int32_t clamp, double *g, double *dg,
                               void *stream);
/* Penalty part of the "node" contact on its own, one pass over every scene's boundary list
 * (bnd_start: B + 1 offsets into bnd; excl_bits / excl_bits_off: the scenes' bit-packed exclusion,
 * bit i * n_be + j of scene e's words excludes its boundary nodes i and j).
 * grad += h2 g, diag += h2 diag (each clamped at 0 when clamp) of every boundary node's own
 * visits. Replaces grad_diag_self_bp_kernel (implicit/kernels.py:772-808) as the adjoint's
 * preconditioner calls it (implicit/adjoint.py:319-336). */
int d2f_node_grad_diag_f64(const double *x, const int32_t *bnd, const int32_t *bnd_start, int32_t B,
                           const uint32_t *excl_bits, const int64_t *excl_bits_off, double radius,
                           double stiffness, double h2, int32_t clamp, double *g, double *dg,
                           void *stream);
/* y += h2 H_node v at x, own rows, unclamped (the contact set is the pairs closer than radius at
 * x). Replaces self_bp_hvp_kernel (implicit/adjoint.py:115-143). */
int d2f_node_hvp_f64(const double *x, const double *v, const int32_t *bnd, const int32_t *bnd_start,
                     int32_t B, const uint32_t *excl_bits, const int64_t *excl_bits_off, double radius,
                     double stiffness, double h2, double *y, void *stream);
/* out[t] += -(1 - Emin) h2 lam . f_t(x) with f_t the element's full-density elastic force (mu0,
 * lam0). Replaces drho_kernel (implicit/adjoint.py:165-178). */
int d2f_elastic_drho_f64(const double *x, const double *lam, const int32_t *tri, const double *Bm,
                         const double *area, const double *mu0, const double *lam0, int32_t T,
                         double h2, double Emin, double *out, void *stream);

/* ---- adjoint of pressure chambers (float64; beyond the reference, whose adjoint refuses
 * pressure: implicit/adjoint.py:376-385; the pressure term is implicit/actuation.py:126-151) --- */
/* y += h2 H_pressure v: each chamber edge (pe_a -> pe_b) adds -1/2 h2 p J(v_b - v_a) to both
 * ends, J(u) = (u_y, -u_x), p = ch_p[pe_ch[k]] (the gradient's derivative; exact, it is linear). */
int d2f_pressure_hvp_f64(const double *v, const int32_t *pe_a, const int32_t *pe_b, const int32_t *pe_ch,
                         const double *ch_p, int32_t n_pe, double h2, double *y, void *stream);
/* out[c] += -lam . dR/dp_c = 1/2 h2 sum over chamber c's edges (lam_a + lam_b) . J(x_b - x_a). */
int d2f_pressure_dp_f64(const double *x, const double *lam, const int32_t *pe_a, const int32_t *pe_b,
                        const int32_t *pe_ch, int32_t n_pe, double h2, double *out, void *stream);

/* ---- adjoint of IPC contact between bodies (float64; beyond the reference, whose adjoint
 * refuses IPCContact: implicit/adjoint.py:376-385; the barrier terms are those of
 * ipc_pair_kernel, implicit/ipc_contact.py:131-296). pairs: the list d2f_ipc_pairs writes at the
 * adjoint's state x (pairs[2 + 2i] edge, pairs[3 + 2i] node id, i < n_pairs); the terms of each
 * pair (interior foot, end corrections with ca / cb) are selected at x and held fixed. */
/* y += h2 (Hessian of the barrier energy) v, UNCLAMPED, friction left out (its tangent is absent,
 * as for the colliders in the reference's adjoint). */
int d2f_ipc_hvp_f64(const double *x, const double *v, const int32_t *pairs, int32_t n_pairs,
                    const int32_t *ea, const int32_t *eb, const double *ca, const double *cb,
                    double dhat, double kappa, double h2, double *y, void *stream);
/* g += h2 barrier gradient, dg += h2 barrier curvature of every term (friction off, nothing
 * masked): the contact part of the adjoint's Jacobi preconditioner (the gradient pass of
 * ipc_pair_kernel mode 0, implicit/ipc_contact.py:131-171). */
int d2f_ipc_grad_diag_f64(const double *x, const int32_t *pairs, int32_t n_pairs, const int32_t *ea,
                          const int32_t *eb, const double *ca, const double *cb, double dhat,
                          double kappa, double h2, double *g, double *dg, void *stream);

/* ---- adjoint with affine bodies (beyond the reference, whose adjoint refuses AffineBodies;
 * the reduction is that of implicit/affine.py:157-224). Slaved node i (slave_of[i] = j >= 0, an
 * index into the scene's an_* lists) follows x_i = J_i q with w_j = an_w[3j .. 3j+2]; the solve
 * runs in reduced vectors (2N doubles): soft nodes' entries in place, group g's
 * q = (t_x, A11, A12, t_y, A21, A22) in the (x, y) slots of nodes slot[3g], slot[3g+1], slot[3g+2]
 * (three of the group's own nodes), every other slaved entry zero. */
/* vf = P vr: soft nodes copied, slaved nodes J_i q. vr and vf must differ. */
int d2f_affine_expand_f64(const double *vr, const int32_t *slave_of, const int32_t *an_grp,
                          const double *an_w, const int32_t *slot, int32_t N, double *vf, void *stream);
/* yr = P^T u with u = yf + mass vf on the slaved nodes when vf is given (the inertia the Krylov
 * solve's diagonal mass leaves out there), or, with square, the diagonal of P^T diag(yf) P.
 * an_first: n_groups + 1 offsets of each group's entries in an_node / an_w. yf and yr must differ. */
int d2f_affine_reduce_f64(const double *yf, const double *vf, const double *mass, const int32_t *slave_of,
                          const int32_t *an_node, const double *an_w, const int32_t *an_first,
                          const int32_t *slot, int32_t N, int32_t n_groups, int32_t square, double *yr,
                          void *stream);

/* ---- MINRES with per-scene Krylov scalars, on the device (replaces the host scipy MINRES of
 * AdjointPNCG2D.adjoint_solve, implicit/adjoint.py:339-369). Scene e solves for the degrees of
 * freedom [2 node_start[e], 2 node_start[e+1]) with its own Lanczos process. All vectors are
 * 2N float64, caller-owned workspace. Protocol: d2f_minres_begin_f64 (x = 0, v = the first
 * Lanczos vector, Kv = 0); then per iteration the caller adds h2 K v into Kv with the
 * Hessian-vector entry points (or writes it), and calls d2f_minres_step_f64, which forms
 * H v = fmask (mass v + Kv), advances every scene still running and leaves that scene's Kv
 * zero for the next product (a finished scene's Kv is left as it is). status[2e] is 0 running,
 * 1 converged (preconditioned residual <= rtol times the initial one), 2 iteration limit,
 * -1 breakdown (a negative preconditioned norm); status[2e+1] its iterations. */
typedef struct d2f_minres {
  int32_t B, maxiter;
  const int32_t *node_start;          /* B+1 */
  const double *mass;                 /* N */
  const double *fmask, *dinv, *b;     /* 2N: 1/0 free mask, Jacobi preconditioner (0 off free), rhs */
  double *x, *v, *r1, *r2, *w1, *w2;  /* 2N each: solution, Lanczos vector, recurrences */
  double *Kv;                         /* 2N: h2 K v on entry to a step; zero on return */
  double *scal;                       /* 8B Krylov scalars */
  int32_t *status;                    /* 2B */
  double rtol;
} d2f_minres;
int d2f_minres_begin_f64(const d2f_minres *m, void *stream);
int d2f_minres_step_f64(const d2f_minres *m, void *stream);

/* ---- allocation-free step -----------------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). d2f_step_scene with
 * the scene kernel's node work arrays (xn, xhat, grad, grad_prev, p, diag, diag_el, x0, bc_q)
 * taken from a static device array of the library instead of the scene's pointers, which are
 * neither read nor written (they may be NULL). Everything else — inputs, x and v, the
 * per-scene scalars, the affine state — is the caller's, as for d2f_step_scene. The array holds
 * D2F_NOALLOC_MAX_NODES nodes (9 x 16 bytes per node: 18 MiB), on every device the library runs
 * on, until the process ends; a larger scene returns D2F_EINVAL. The first call on a device looks
 * the array up (hipGetSymbolAddress); no call allocates. Every call on one device shares that
 * device's array: calls that overlap on one device (two streams or two threads) race. */
#define D2F_NOALLOC_MAX_NODES 131072
int d2f_step_scene_noalloc(const d2f_scene *s, const d2f_params *p, void *stream);

#ifdef __cplusplus
}
#endif
#endif
