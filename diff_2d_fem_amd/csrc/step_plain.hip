// step_plain.hip — the per-kernel float64 path: one launch per reference kernel family, in
// the order of PNCG2DSolver._iterate (reference implicit/solver.py:252-451), with atomics for
// every scatter and per-environment reduction. It is the plain GPU mirror of the reference's
// launch structure and the general path for scenes of any size.
#include <stdarg.h>
#include <string.h>

#include <vector>

#include <climits>

#include "common.h"
#include "node_contact.h"

namespace d2f {

static thread_local char g_err[512];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

using S = d2f_scene;
using Pm = d2f_params;

constexpr int TPB = 256;

#define GRID_LOOP(i, n)                                                                 \
  for (int base_ = blockIdx.x * blockDim.x; base_ < (n); base_ += gridDim.x * blockDim.x) \
    for (int i = base_ + threadIdx.x, once_ = 1; once_; once_ = 0)

__device__ __forceinline__ double fr_of(const S &s, int i) { return s.free_[i] ? 1.0 : 0.0; }
// scene e is stepped elsewhere (the batch's routing): leave its state alone
__device__ __forceinline__ bool skip_env(const S &s, int e) { return s.env_skip && s.env_skip[e]; }
__device__ __forceinline__ bool skip_node(const S &s, int i) { return s.env_skip && s.env_skip[s.node2env[i]]; }

// ------------------------------------------------------------------------- step setup
// reference implicit/kernels.py:42-68, 438-441
__global__ void k_init_step(S s, double dt, double drag) {
  GRID_LOOP(i, s.N) {
    if (i < s.N && !skip_node(s, i)) {
      const V2<double> x = ld2(s.x, i);
      st2(s.xn, i, x);
      st2(s.xhat, i, x + (dt * drag) * ld2(s.v, i));
    }
  }
}
__global__ void k_apply_bc(S s) {
  GRID_LOOP(b, s.n_bc) {
    if (b < s.n_bc && !skip_node(s, s.bc_idx[b])) {
      const int n = s.bc_idx[b];
      const V2<double> t = ld2(s.bc_target, b);
      st2(s.x, n, t);
      st2(s.xhat, n, t);
    }
  }
}
__global__ void k_reset_step(S s) {
  GRID_LOOP(e, s.B) {
    if (e < s.B) {
      s.dE0[e] = 0.0;
      if (skip_env(s, e)) {
        s.active[e] = 0;     // its iteration count stays as its own path left it
      } else {
        s.active[e] = 1;
        s.iters[e] = 0;
      }
    }
  }
}
__global__ void k_commit_v(S s, double inv_dt) {
  GRID_LOOP(i, s.N) {
    if (i < s.N && !skip_node(s, i)) st2(s.v, i, inv_dt * (ld2(s.x, i) - ld2(s.xn, i)));
  }
}

// ------------------------------------------------------------------------- gradient + diagonal
// reference implicit/kernels.py:73-82 (+ the external force, implicit/actuation.py:37-41)
// Also clears the direction pass's per-scene dot-product accumulators (k_mask_dk and the mode
// reductions add into them later in the same iteration; N >= B).
__global__ void k_grad_diag_mass(S s, double h2) {
  GRID_LOOP(i, s.N) {
    if (i < s.B) {
      s.gPg[i] = 0.0;
      s.g_p[i] = 0.0;
      s.g_Py[i] = 0.0;
      s.y_p[i] = 0.0;
      s.y_Py[i] = 0.0;
    }
    if (i < s.N) {
      const V2<double> g = ld2(s.grad, i);
      st2(s.grad_prev, i, g);
      const double m = s.mass[i], fr = fr_of(s, i);
      const V2<double> d = ld2(s.x, i) - ld2(s.xhat, i);
      st2(s.grad, i, V2<double>{m * d.x * fr, m * d.y * fr});
      st2(s.diag, i, V2<double>{m, m});
    }
  }
}

template <typename R>
__device__ __forceinline__ M2<R> load_B(const double *Bm, int t) {
  const double4 b = reinterpret_cast<const double4 *>(Bm)[t];
  return {R(b.x), R(b.y), R(b.z), R(b.w)};
}

// Elastic gradient and diagonal (reference implicit/kernels.py:86-128), tiled: workgroup c
// takes triangles [c * D2F_ELEM_CHUNK, (c + 1) * D2F_ELEM_CHUNK), ELEM_PER_THREAD per thread.
// It gathers the positions of the chunk's nodes (the tile's node list) into LDS, sums every
// triangle's contributions per node in LDS (the triangles address the list through their
// 10-bit local indices), then adds each node it owns once (plain read-modify-write: no other
// workgroup touches it). Its sums at halo nodes (shared with other chunks):
// * PARTIAL: written to their partial slots, which k_elastic_halo adds per node in chunk order.
//   Used when the launch's working set exceeds the Infinity Cache: there the device-scope
//   float64 atomics on the halo nodes took about a quarter of the launch (46 M triangles:
//   1.46 -> 1.37 ms).
// * otherwise: device-scope atomics, which resolve in cache at these sizes and save the second
//   launch (0.9 M triangles: 19.9 against 23.2 us).
constexpr int ELEM_TPB = 256;   // 4 triangles per thread: 4 waves per workgroup
constexpr int ELEM_PER_THREAD = D2F_ELEM_CHUNK / ELEM_TPB;
constexpr int ELEM_PARTIAL_MIN_T = 1 << 22;   // 4.2 M triangles: 285 MB of float64 constants

template <bool PARTIAL>
__global__ __launch_bounds__(ELEM_TPB) void k_elastic_grad_diag_tiled(
    const double *x, const double *Bm, const double *area, const double *mu, const double *lam, int T,
    double h2, int clamp, d2f_elem_tiles tl, double *grad, double *diag) {
  // nodes of the chunk, one plane per component (max_nodes each): x, y, grad x, grad y, diag x,
  // diag y; neighbouring lanes then touch neighbouring banks
  extern __shared__ double sh[];
  // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous run of chunks
  // so that the nodes neighbouring chunks share stay in one L2.
  const int nbk = gridDim.x, q8 = nbk / 8, r8 = nbk % 8, xcd = blockIdx.x % 8;
  const int c = xcd * q8 + min(xcd, r8) + blockIdx.x / 8;
  const int n0 = tl.node_off[c], nn = tl.node_off[c + 1] - n0;
  const int h0 = tl.halo_off[c], no = nn - (tl.halo_off[c + 1] - h0);   // owned entries first
  const int M = tl.max_nodes;
  double *xs = sh, *ys = sh + M, *acc = sh + 2 * M;
  for (int j = threadIdx.x; j < nn; j += ELEM_TPB) {
    const V2<double> p = ld2(x, tl.nodes[n0 + j]);
    xs[j] = p.x;
    ys[j] = p.y;
    acc[j] = acc[M + j] = acc[2 * M + j] = acc[3 * M + j] = 0.0;
  }
  __syncthreads();
  const int base = c * D2F_ELEM_CHUNK + threadIdx.x;
#pragma unroll
  for (int k = 0; k < ELEM_PER_THREAD; ++k) {
    const int t = base + k * ELEM_TPB;
    if (t >= T) continue;
    const uint32_t l = tl.tri_local[t];
    const int li[3] = {int(l & 1023u), int((l >> 10) & 1023u), int((l >> 20) & 1023u)};
    const V2<double> p0 = {xs[li[0]], ys[li[0]]}, p1 = {xs[li[1]], ys[li[1]]}, p2 = {xs[li[2]], ys[li[2]]};
    const M2<double> B = load_B<double>(Bm, t);
    const double m_ = mu[t], l_ = lam[t], A = area[t];
    const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
    const NH<double> h = nh_eval(mat_mul(Ds, B), m_, l_);
    V2<double> g[3], d[3];
    nh_grad(h, B, m_, A, g[0], g[1], g[2]);
    nh_diag(h, B, m_, l_, A, d[0], d[1], d[2]);
    if (clamp)
#pragma unroll
      for (int v = 0; v < 3; ++v) d[v] = {fmax(d[v].x, 0.0), fmax(d[v].y, 0.0)};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      double *a = acc + li[v];
      atomicAdd(a, g[v].x);
      atomicAdd(a + M, g[v].y);
      atomicAdd(a + 2 * M, d[v].x);
      atomicAdd(a + 3 * M, d[v].y);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nn; j += ELEM_TPB) {
    const double *a = acc + j;
    const V2<double> gg = {h2 * a[0], h2 * a[M]}, dd = {h2 * a[2 * M], h2 * a[3 * M]};
    if (j < no) {
      const int n = tl.nodes[n0 + j];
      st2(grad, n, ld2(grad, n) + gg);
      st2(diag, n, ld2(diag, n) + dd);
    } else if (PARTIAL) {
      reinterpret_cast<double4 *>(tl.part)[tl.hdest[h0 + j - no]] = make_double4(gg.x, gg.y, dd.x, dd.y);
    } else {
      const int n = tl.nodes[n0 + j];
      atomic_add2(grad, n, gg);
      atomic_add2(diag, n, dd);
    }
  }
}

// The halo nodes of the tiled pass: node hnode[k] adds its chunks' partial sums (contiguous,
// in chunk order). diag null: the partials carry only the first two components (the
// Hessian-vector product).
__global__ void k_elastic_halo(d2f_elem_tiles tl, double *grad, double *diag) {
  GRID_LOOP(k, tl.n_hnodes) {
    if (k < tl.n_hnodes) {
      const int n = tl.hnode[k];
      V2<double> g = ld2(grad, n), d = {0.0, 0.0};
      if (diag) d = ld2(diag, n);
      for (int i = tl.hnode_off[k]; i < tl.hnode_off[k + 1]; ++i) {
        const double4 q = reinterpret_cast<const double4 *>(tl.part)[i];
        g = g + V2<double>{q.x, q.y};
        d = d + V2<double>{q.z, q.w};
      }
      st2(grad, n, g);
      if (diag) st2(diag, n, d);
    }
  }
}

// Pressure chamber edges handed to the tiled Hessian-vector product: chunk c takes the edges
// pe_off[c] .. pe_off[c + 1], whose two nodes are entries lo | hi << 10 of its node list, with
// their chamber's index into ch_p. pe_off null: no pressure.
struct PressureTiles {
  const int32_t *pe_off;
  const uint32_t *pe_loc;
  const int32_t *pe_ch;
  const double *ch_p;
};

// Elastic Hessian-vector product y += h2 (d2Psi/dx2) v (reference implicit/adjoint.py:70-85),
// tiled as k_elastic_grad_diag_tiled: chunk c gathers its nodes' positions and direction into
// LDS, sums every triangle's product per node there, then adds each owned node once (plain
// read-modify-write) and its halo sums by device-scope atomics, or (PARTIAL) through the
// tiles' partial slots and k_elastic_halo. With pressure edges, their product (k_pressure_hvp's
// term) is summed into the same LDS sums.
template <bool PARTIAL>
__global__ __launch_bounds__(ELEM_TPB) void k_elastic_hvp_tiled(const double *x, const double *vv,
                                                                const double *Bm, const double *area,
                                                                const double *mu, const double *lam, int T,
                                                                double h2, d2f_elem_tiles tl, PressureTiles pt,
                                                                double *y) {
  // planes of max_nodes each: x, y, v_x, v_y, (Hv)_x, (Hv)_y
  extern __shared__ double sh[];
  const int nbk = gridDim.x, q8 = nbk / 8, r8 = nbk % 8, xcd = blockIdx.x % 8;
  const int c = xcd * q8 + min(xcd, r8) + blockIdx.x / 8;
  const int n0 = tl.node_off[c], nn = tl.node_off[c + 1] - n0;
  const int h0 = tl.halo_off[c], no = nn - (tl.halo_off[c + 1] - h0);
  const int M = tl.max_nodes;
  double *xs = sh, *ys = sh + M, *vxs = sh + 2 * M, *vys = sh + 3 * M, *acc = sh + 4 * M;
  for (int j = threadIdx.x; j < nn; j += ELEM_TPB) {
    const int n = tl.nodes[n0 + j];
    const V2<double> p = ld2(x, n), q = ld2(vv, n);
    xs[j] = p.x;
    ys[j] = p.y;
    vxs[j] = q.x;
    vys[j] = q.y;
    acc[j] = acc[M + j] = 0.0;
  }
  __syncthreads();
  const int base = c * D2F_ELEM_CHUNK + threadIdx.x;
#pragma unroll
  for (int k = 0; k < ELEM_PER_THREAD; ++k) {
    const int t = base + k * ELEM_TPB;
    if (t >= T) continue;
    const uint32_t l = tl.tri_local[t];
    const int li[3] = {int(l & 1023u), int((l >> 10) & 1023u), int((l >> 20) & 1023u)};
    const V2<double> p0 = {xs[li[0]], ys[li[0]]}, p1 = {xs[li[1]], ys[li[1]]}, p2 = {xs[li[2]], ys[li[2]]};
    const M2<double> B = load_B<double>(Bm, t);
    const double m_ = mu[t], l_ = lam[t], A = area[t];
    const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
    const NH<double> h = nh_eval(mat_mul(Ds, B), m_, l_);
    const M2<double> dF = dF_of(V2<double>{vxs[li[0]], vys[li[0]]}, V2<double>{vxs[li[1]], vys[li[1]]},
                                V2<double>{vxs[li[2]], vys[li[2]]}, B);
    const M2<double> dP = nh_d2_dir(h, dF, m_, l_);
    const V2<double> f1 = {A * (dP.a * B.a + dP.b * B.b), A * (dP.c * B.a + dP.d * B.b)};
    const V2<double> f2 = {A * (dP.a * B.c + dP.b * B.d), A * (dP.c * B.c + dP.d * B.d)};
    const V2<double> f[3] = {{-(f1.x + f2.x), -(f1.y + f2.y)}, f1, f2};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      atomicAdd(acc + li[v], f[v].x);
      atomicAdd(acc + M + li[v], f[v].y);
    }
  }
  if (pt.pe_off) {
    // -1/2 p J(v_b - v_a), J(u) = (u_y, -u_x), on both ends (h2 is applied below)
    for (int q = pt.pe_off[c] + threadIdx.x; q < pt.pe_off[c + 1]; q += ELEM_TPB) {
      const uint32_t l = pt.pe_loc[q];
      const int la = int(l & 1023u), lb = int((l >> 10) & 1023u);
      const double cp = -0.5 * pt.ch_p[pt.pe_ch[q]];
      const V2<double> u = {vxs[lb] - vxs[la], vys[lb] - vys[la]};
      const V2<double> w = {cp * u.y, -cp * u.x};
      atomicAdd(acc + la, w.x);
      atomicAdd(acc + M + la, w.y);
      atomicAdd(acc + lb, w.x);
      atomicAdd(acc + M + lb, w.y);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nn; j += ELEM_TPB) {
    const V2<double> yy = {h2 * acc[j], h2 * acc[M + j]};
    if (j < no) {
      const int n = tl.nodes[n0 + j];
      st2(y, n, ld2(y, n) + yy);
    } else if (PARTIAL) {
      reinterpret_cast<double4 *>(tl.part)[tl.hdest[h0 + j - no]] = make_double4(yy.x, yy.y, 0.0, 0.0);
    } else {
      atomic_add2(y, tl.nodes[n0 + j], yy);
    }
  }
}

// The same with one atomic per contribution (no tiles, or a chunk too large for LDS).
__global__ void k_elastic_grad_diag(const double *x, const int32_t *tri, const double *Bm,
                                    const double *area, const double *mu, const double *lam, int T,
                                    double h2, int clamp, double *grad, double *diag) {
  GRID_LOOP(t, T) {
    if (t < T) {
      const int ii[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
      const V2<double> p0 = ld2(x, ii[0]), p1 = ld2(x, ii[1]), p2 = ld2(x, ii[2]);
      const M2<double> B = load_B<double>(Bm, t);
      const double m_ = mu[t], l_ = lam[t], A = area[t];
      const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
      const NH<double> h = nh_eval(mat_mul(Ds, B), m_, l_);
      V2<double> g[3], d[3];
      nh_grad(h, B, m_, A, g[0], g[1], g[2]);
      nh_diag(h, B, m_, l_, A, d[0], d[1], d[2]);
      if (clamp)
#pragma unroll
        for (int v = 0; v < 3; ++v) d[v] = {fmax(d[v].x, 0.0), fmax(d[v].y, 0.0)};
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        atomic_add2(grad, ii[v], h2 * g[v]);
        atomic_add2(diag, ii[v], h2 * d[v]);
      }
    }
  }
}

static int launch_elastic_grad_diag(const double *x, const int32_t *tri, const double *Bm,
                                    const double *area, const double *mu, const double *lam, int T,
                                    double h2, int clamp, const d2f_elem_tiles *tiles, double *grad,
                                    double *diag, hipStream_t st) {
  if (T == 0) return 0;
  if (tiles && tiles->node_off && tiles->max_nodes > 0 && tiles->max_nodes <= D2F_ELEM_NODES_MAX) {
    D2F_CHECK(tiles->nodes && tiles->halo_off && tiles->tri_local && tiles->part && tiles->n_hnodes >= 0 &&
                  (tiles->n_hnodes == 0 || (tiles->hnode && tiles->hnode_off && tiles->hdest)),
              "elastic grad/diag: incomplete tiles");
    const int blocks = (T + D2F_ELEM_CHUNK - 1) / D2F_ELEM_CHUNK;
    const size_t lds = sizeof(double) * 6 * (size_t)tiles->max_nodes;   // at most 48 KiB
    const bool partial = T >= ELEM_PARTIAL_MIN_T;
    if (partial)
      hipLaunchKernelGGL(k_elastic_grad_diag_tiled<true>, dim3(blocks), dim3(ELEM_TPB), lds, st, x, Bm, area, mu,
                         lam, T, h2, clamp, *tiles, grad, diag);
    else
      hipLaunchKernelGGL(k_elastic_grad_diag_tiled<false>, dim3(blocks), dim3(ELEM_TPB), lds, st, x, Bm, area, mu,
                         lam, T, h2, clamp, *tiles, grad, diag);
    if (partial && tiles->n_hnodes > 0) {
      const int hb = (int)std::min<long>((tiles->n_hnodes + 255) / 256, 65536L);
      hipLaunchKernelGGL(k_elastic_halo, dim3(hb), dim3(256), 0, st, *tiles, grad, diag);
    }
  } else {
    const int blocks = (int)std::min<long>((T + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_elastic_grad_diag, dim3(blocks), dim3(256), 0, st, x, tri, Bm, area, mu, lam, T, h2,
                       clamp, grad, diag);
  }
  D2F_HIP(hipGetLastError());
  return 0;
}

// Collider barrier + lagged friction of every boundary node (one thread owns its node, so no
// atomics): core/sdf.py:141-154, 213-233; implicit/kernels.py:132-157, 661-690.
__global__ void k_colliders_grad_diag(S s, Pm p) {
  GRID_LOOP(b, s.n_b) {
    if (b < s.n_b) {
      const int n = s.bnd[b];
      const V2<double> x = ld2(s.x, n);
      V2<double> g = ld2(s.grad, n), dg = ld2(s.diag, n);
      for (int k = 0; k < s.n_col; ++k) {
        V2<double> gg, dd;
        if (collider_grad_diag(to_collider(s.col[k]), x, p.dhat, p.kappa, p.h2, 1, gg, dd)) {
          g = g + gg;
          dg = dg + dd;
        }
      }
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          const double sc = p.h2 * c.mu * f.f_n / fmax(f.vtm, p.epsv);
          g = g + sc * f.vt;
          const double kt = p.h2 * c.mu * f.f_n * p.inv_dt / fmax(f.vtm, p.epsv);
          const V2<double> th = {-f.n.y, f.n.x};
          dg = dg + V2<double>{kt * th.x * th.x, kt * th.y * th.y};
        }
      }
      st2(s.grad, n, g);
      st2(s.diag, n, dg);
    }
  }
}

// grad -= h2 f (implicit/actuation.py:37-41)
__global__ void k_extforce_grad(S s, double h2) {
  GRID_LOOP(i, s.N) {
    if (i < s.N) st2(s.grad, i, ld2(s.grad, i) - h2 * ld2(s.fext, i));
  }
}

// ------------------------------------------------------------------------- IPC pairs
struct PairCtx {
  double dhat, kappa, mu, epsv, inv_dt, h2, reach, gap, floor_;
  int per_body;
  int *plist;  // mode 4: plist[0] counts the pairs, (edge, node) pairs from plist[2] on
  int pcap;    // capacity of plist in pairs
};

// The scene's boundary nodes as a pair pass reads them: positions, previous positions and the
// pass's direction from LDS copies indexed by local boundary index when staged (x, xn, pd not
// null), else from global memory by node id; exclusion from an LDS bitset when staged (bits
// not null: bit jp * n_be + j excludes local nodes jp and j), else from the byte matrix.
struct BView {
  const double *x, *xn, *pd;
  const uint32_t *bits;
};

// One barrier-and-friction term (reference implicit/ipc_contact.py:131-171) of node ip against
// the edge point wa x_a + wb x_b. sp, sa, sb: the three nodes' slips x - xn; qp, qa, qb: their
// direction (mode 1). Mode 0 adds into grad / diag; modes 1 (pHp) and 2 (energy) add the
// term's scalar to acc, which the caller reduces into the environment's total.
__device__ __forceinline__ void ipc_term(int mode, const PairCtx &c, double coef, int ip, int ia, int ib,
                                         double wa, double wb, double d, V2<double> n, double ddd, int e,
                                         double mu, V2<double> sp, V2<double> sa, V2<double> sb,
                                         V2<double> qp, V2<double> qa, V2<double> qb, double *grad,
                                         double *diag, const uint8_t *freev, double &acc) {
  const double bp = barrier_bp(d, c.dhat);
  const double f_n = -c.kappa * bp;
  const V2<double> tn = {-n.y, n.x};
  const V2<double> u = sp - wa * sa - wb * sb;
  const double ut = dot(u, tn);
  const double vt = fabs(ut) * c.inv_dt;
  const double kt = mu * f_n * c.inv_dt / fmax(vt, c.epsv);
  if (mode == 0) {
    const double bpp = barrier_bpp(d, c.dhat);
    const V2<double> g = c.h2 * ((-f_n) * n + (mu * f_n * (ut * c.inv_dt / fmax(vt, c.epsv))) * tn);
    const V2<double> hd = {c.h2 * (c.kappa * bpp * n.x * n.x + kt * tn.x * tn.x),
                           c.h2 * (c.kappa * bpp * n.y * n.y + kt * tn.y * tn.y)};
    const double fp = freev[ip] ? 1.0 : 0.0, fa = freev[ia] ? 1.0 : 0.0, fb = freev[ib] ? 1.0 : 0.0;
    atomic_add2(grad, ip, (coef * fp) * g);
    atomic_add2(diag, ip, coef * hd);
    atomic_add2(grad, ia, (-coef * wa * fa) * g);
    atomic_add2(diag, ia, (coef * wa * wa) * hd);
    atomic_add2(grad, ib, (-coef * wb * fb) * g);
    atomic_add2(diag, ib, (coef * wb * wb) * hd);
  } else if (mode == 1) {
    const double bpp = barrier_bpp(d, c.dhat);
    const V2<double> dq = qp - wa * qa - wb * qb;
    const double dn = dot(dq, n), dt_ = dot(dq, tn);
    const double val = fmax(c.kappa * (bpp * dn * dn + bp * ddd), 0.0);
    acc += coef * c.h2 * (val + kt * dt_ * dt_);
  } else {
    const double asl = c.epsv / c.inv_dt;
    const double au = fabs(ut);
    const double f0 = au < asl ? au * au / (2.0 * asl) : au - 0.5 * asl;
    acc += coef * c.h2 * (c.kappa * barrier_b(d, c.dhat) + mu * f_n * f0);
  }
}

// One boundary edge as the pair terms read it (positions, slips x - xn and direction of its
// ends as the mode needs them).
struct EdgeData {
  int a, b, e, la, lb;
  V2<double> xa, xb, sa, sb, qa, qb, ab;
  double L2, L, ca, cb;
};

__device__ __forceinline__ V2<double> bv_x(const S &s, const BView &bv, int jl, int node) {
  return bv.x ? V2<double>{bv.x[2 * jl], bv.x[2 * jl + 1]} : ld2(s.x, node);
}
__device__ __forceinline__ V2<double> bv_xn(const S &s, const BView &bv, int jl, int node) {
  return bv.xn ? V2<double>{bv.xn[2 * jl], bv.xn[2 * jl + 1]} : ld2(s.xn, node);
}
__device__ __forceinline__ V2<double> bv_pd(const BView &bv, const double *pdir, int jl, int node) {
  return bv.pd ? V2<double>{bv.pd[2 * jl], bv.pd[2 * jl + 1]} : ld2(pdir, node);
}

__device__ __forceinline__ EdgeData edge_data(const S &s, int k, int mode, const BView &bv, const double *pdir) {
  EdgeData d;
  d.a = s.ea[k];
  d.b = s.eb[k];
  d.e = s.e2env[k];
  d.la = s.ela[k];
  d.lb = s.elb[k];
  d.ca = s.ca[k];
  d.cb = s.cb[k];
  d.xa = bv_x(s, bv, d.la, d.a);
  d.xb = bv_x(s, bv, d.lb, d.b);
  const bool slips = mode == 0 || mode == 1 || mode == 2;
  const bool dirs = mode == 1 || mode == 3;
  d.sa = slips ? d.xa - bv_xn(s, bv, d.la, d.a) : V2<double>{0.0, 0.0};
  d.sb = slips ? d.xb - bv_xn(s, bv, d.lb, d.b) : V2<double>{0.0, 0.0};
  d.qa = dirs ? bv_pd(bv, pdir, d.la, d.a) : V2<double>{0.0, 0.0};
  d.qb = dirs ? bv_pd(bv, pdir, d.lb, d.b) : V2<double>{0.0, 0.0};
  d.ab = d.xb - d.xa;
  d.L2 = fmax(dot(d.ab, d.ab), 1.0e-30);
  d.L = sqrt(d.L2);
  return d;
}

// The pair filter (reference implicit/ipc_contact.py:212-235): node jl of the edge's scene is
// not an end, is closer than r to the edge, and is not excluded against either end.
__device__ __forceinline__ bool pair_in(const S &s, const EdgeData &ed, int jl, V2<double> xp, double r,
                                        int n_be, int64_t eoff, const BView &bv) {
  if (jl == ed.la || jl == ed.lb) return false;
  const V2<double> w = xp - ed.xa;
  const double tt = dot(w, ed.ab) / ed.L2;
  V2<double> dvec = w - tt * ed.ab;
  if (tt <= 0.0) dvec = w;
  else if (tt >= 1.0) dvec = xp - ed.xb;
  if (!(len(dvec) < r)) return false;
  if (bv.bits) {
    const int64_t i0 = (int64_t)jl * n_be + ed.la, i1 = (int64_t)jl * n_be + ed.lb;
    return !(((bv.bits[i0 >> 5] >> (i0 & 31)) & 1u) || ((bv.bits[i1 >> 5] >> (i1 & 31)) & 1u));
  }
  const uint32_t *xw = s.excl_bits + eoff;
  return !(bit_of(xw, (int64_t)jl * n_be + ed.la) || bit_of(xw, (int64_t)jl * n_be + ed.lb));
}

// The terms of one pair that passed the filter within the reach (reference
// implicit/ipc_contact.py:175-296): clearance and step limit (modes 1, 3), then the interior
// term and the endpoint terms with their vertex-correction coefficients. amin / cmin take the
// minimum with the pair's step limit and clearance, acc adds its pHp (mode 1) or energy
// (mode 2). ip: the node's id (-1: read it from the scene's boundary list).
__device__ __forceinline__ void pair_body(const S &s, int k, int mode, const PairCtx &c, const EdgeData &ed,
                                          int jl, int ip, V2<double> xp, const double *pdir, double *grad,
                                          double *diag, const uint8_t *freev, double t_c,
                                          const double *mu_node, const BView &bv, double &amin,
                                          double &cmin, double &acc) {
  const bool dirs = mode == 1 || mode == 3;
  if (ip < 0) ip = s.bnd[s.bnd_start[ed.e] + jl];
  if (mode == 4) {
    const int slot = atomicAdd(c.plist, 1);
    if (slot < c.pcap) {
      c.plist[2 + 2 * slot] = k;
      c.plist[3 + 2 * slot] = ip;
    }
    return;
  }
  const double dlo = 1.0e-9 * c.dhat, onset = 0.02 * c.dhat;
  const V2<double> w = xp - ed.xa;
  const double tt = dot(w, ed.ab) / ed.L2;
  V2<double> dvec = w - tt * ed.ab;
  if (tt <= 0.0) dvec = w;
  else if (tt >= 1.0) dvec = xp - ed.xb;
  double d = len(dvec);
  const V2<double> qp = dirs ? bv_pd(bv, pdir, jl, ip) : V2<double>{0.0, 0.0};
  if (dirs) {
    cmin = fmin(cmin, d);
    double keep = fmax(c.gap * d, c.dhat - onset);
    if (d < c.dhat) keep = d - fmin(2.0 * (c.dhat - d) + onset, 0.5 * d);
    if (mode == 3) keep = fmin(d, fmax(c.gap * d, c.floor_));
    amin = fmin(amin, accd(xp, ed.xa, ed.xb, qp, ed.qa, ed.qb, keep, t_c));
  }
  if (mode == 3) return;
  if (!(d < c.dhat || len(w) < c.dhat || len(xp - ed.xb) < c.dhat)) return;
  const V2<double> sp = xp - bv_xn(s, bv, jl, ip);
  double mu_p = c.mu;
  if (c.per_body) {
    const double mi = mu_node[ip], mj = mu_node[ed.a], sm = mi + mj;
    mu_p = sm > 0.0 ? 2.0 * mi * mj / sm : 0.0;
  }
  const int a = ed.a, b = ed.b, e = ed.e;
  if (tt > 0.0 && tt < 1.0 && d < c.dhat) {
    d = fmax(d, dlo);
    double ddd = 0.0;
    if (mode == 1) {
      const V2<double> de = ed.qb - ed.qa;
      const V2<double> dw = qp - ed.qa;
      const double c0 = cross(ed.ab, w);
      const double c1 = cross(de, w) + cross(ed.ab, dw);
      const double c2 = 2.0 * cross(de, dw);
      const double L1 = dot(ed.ab, de) / ed.L;
      const double Lq = (dot(de, de) - L1 * L1) / ed.L;
      const double s0 = c0 / ed.L;
      const double s1 = (c1 - s0 * L1) / ed.L;
      ddd = (c2 - 2.0 * s1 * L1 - s0 * Lq) / ed.L;
      if (c0 < 0.0) ddd = -ddd;
    }
    ipc_term(mode, c, 1.0, ip, a, b, 1.0 - tt, tt, d, (1.0 / d) * dvec, ddd, e, mu_p, sp, ed.sa, ed.sb, qp,
             ed.qa, ed.qb, grad, diag, freev, acc);
  }
  double coef_a = -ed.ca;
  if (tt <= 0.0) coef_a += 1.0;
  double d_a = len(w);
  if (coef_a != 0.0 && d_a < c.dhat) {
    d_a = fmax(d_a, dlo);
    const V2<double> n_a = (1.0 / d_a) * w;
    double ddd = 0.0;
    if (mode == 1) {
      const V2<double> dw = qp - ed.qa;
      const double dn = dot(dw, n_a);
      ddd = (dot(dw, dw) - dn * dn) / d_a;
    }
    ipc_term(mode, c, coef_a, ip, a, b, 1.0, 0.0, d_a, n_a, ddd, e, mu_p, sp, ed.sa, ed.sb, qp, ed.qa, ed.qb,
             grad, diag, freev, acc);
  }
  double coef_b = -ed.cb;
  if (tt >= 1.0) coef_b += 1.0;
  const V2<double> wb_ = xp - ed.xb;
  double d_b = len(wb_);
  if (coef_b != 0.0 && d_b < c.dhat) {
    d_b = fmax(d_b, dlo);
    const V2<double> n_b = (1.0 / d_b) * wb_;
    double ddd = 0.0;
    if (mode == 1) {
      const V2<double> dw = qp - ed.qb;
      const double dn = dot(dw, n_b);
      ddd = (dot(dw, dw) - dn * dn) / d_b;
    }
    ipc_term(mode, c, coef_b, ip, a, b, 0.0, 1.0, d_b, n_b, ddd, e, mu_p, sp, ed.sa, ed.sb, qp, ed.qa, ed.qb,
             grad, diag, freev, acc);
  }
}

// t_c (the step the ACCD may stop at) is each environment's alpha at launch, read from
// a snapshot so that every pair sees the same value. One thread per edge, every boundary node of
// its scene a candidate.
__global__ void k_ipc(S s, int mode, PairCtx c, const double *pdir, double *grad, double *diag,
                      const uint8_t *freev, double *E, double *clr, double *alpha, const double *tc) {
  GRID_LOOP(k, s.n_e) {
    if (k < s.n_e) {
      const BView bv{nullptr, nullptr, nullptr, nullptr};
      const EdgeData ed = edge_data(s, k, mode, bv, pdir);
      const int start = s.bnd_start[ed.e], n_be = s.bnd_start[ed.e + 1] - start;
      const int64_t eoff = s.excl_bits_off[ed.e];
      const double t_c = (mode == 1 || mode == 3) ? tc[ed.e] : 0.0;
      double amin = BIG, cmin = BIG, acc = 0.0;
      for (int jl = 0; jl < n_be; ++jl) {
        const int pn = s.bnd[start + jl];
        const V2<double> xp = ld2(s.x, pn);
        if (pair_in(s, ed, jl, xp, c.reach, n_be, eoff, bv))
          pair_body(s, k, mode, c, ed, jl, pn, xp, pdir, grad, diag, freev, t_c, s.mu_node, bv, amin, cmin, acc);
      }
      if (cmin < BIG) atomic_min_d(clr + ed.e, cmin);
      if (amin < BIG) atomic_min_d(alpha + ed.e, amin);
      if (acc != 0.0) atomicAdd((mode == 1 ? s.pHp : E) + ed.e, acc);
    }
  }
}

// The same pass with workgroups per environment. The environment's boundary positions
// (previous positions and direction as the mode needs them, and its exclusion bitset when it
// fits) are staged in LDS, and a uniform grid over them is built there by counting sort (cells
// at least the contact reach wide, at most IPC_CELLS of them): an edge visits the nodes of the
// cells its box, widened by the reach, overlaps (the reference queries a hash grid,
// implicit/ipc_contact.py:380-391). LDS: 8 + 16 per staged vector bytes per boundary node, 4 per
// cell, 4 per bitset word.
constexpr int IPC_CELLS = 2048;
constexpr int IPC_TPB = 256;
constexpr int IPC_TEAM = 8;  // lanes per edge in the candidate search
// workgroups per environment, each building the grid and taking every IPC_PARTS-th round of
// edges: more resident workgroups per CU than one per environment (the kernel's registers
// allow two waves per SIMD)
constexpr int IPC_PARTS = 4;
// the gradient and energy passes keep only pairs within dhat, few: one workgroup per
// environment builds its grid once and works them all
__host__ __device__ constexpr int ipc_parts(int mode) { return (mode == 1 || mode == 3) ? IPC_PARTS : 1; }

// LDS bytes of k_ipc_env for a mode: bitset words excluded
__host__ __device__ inline size_t ipc_env_lds(int mode, int max_nb) {
  const int vecs = 1 + ((mode == 0 || mode == 1 || mode == 2) ? 1 : 0) + ((mode == 1 || mode == 3) ? 1 : 0);
  return (16 * (size_t)vecs + 8) * (size_t)max_nb + sizeof(int) * (IPC_CELLS + 2);
}

template <int mode>
__global__ __launch_bounds__(IPC_TPB) void k_ipc_env(S s, PairCtx c, const double *pdir, double *grad,
                                                     double *diag, const uint8_t *freev, double *E,
                                                     double *clr, double *alpha, const double *tc,
                                                     int bits_words, int plist_cap) {
  extern __shared__ double sh[];
  constexpr bool slips = mode == 0 || mode == 1 || mode == 2;
  constexpr bool dirs = mode == 1 || mode == 3;
  double *bp = sh;
  double *bxn = bp + 2 * s.max_nb;
  double *bpd = bxn + (slips ? 2 * s.max_nb : 0);
  int *citem = reinterpret_cast<int *>(bpd + (dirs ? 2 * s.max_nb : 0));
  int *cstart = citem + s.max_nb;                  // IPC_CELLS + 2
  uint32_t *bits = reinterpret_cast<uint32_t *>(cstart + IPC_CELLS + 2);
  uint32_t *plist = bits + bits_words;             // plist_cap packed pairs
  __shared__ double lo[2], cell;
  __shared__ int ncx, ncy, use_bits;
  __shared__ double red[4 * (IPC_TPB / WAVE)];
  const int e = blockIdx.x;
  const int start = s.bnd_start[e], n_be = s.bnd_start[e + 1] - start;
  if (s.e_start[e + 1] == s.e_start[e]) return;    // no edges in this environment
  double mnx = BIG, mny = BIG, mxx = -BIG, mxy = -BIG;
  for (int j = threadIdx.x; j < n_be; j += IPC_TPB) {
    const int n = s.bnd[start + j];
    const V2<double> x = ld2(s.x, n);
    bp[2 * j] = x.x;
    bp[2 * j + 1] = x.y;
    if (slips) {
      const V2<double> q = ld2(s.xn, n);
      bxn[2 * j] = q.x;
      bxn[2 * j + 1] = q.y;
    }
    if (dirs) {
      const V2<double> q = ld2(pdir, n);
      bpd[2 * j] = q.x;
      bpd[2 * j + 1] = q.y;
    }
    mnx = fmin(mnx, x.x);
    mny = fmin(mny, x.y);
    mxx = fmax(mxx, x.x);
    mxy = fmax(mxy, x.y);
  }
  // the exclusion bitset, when this scene's fits the words the launch reserved
  const int64_t w0 = (bits_words > 0 && s.excl_bits) ? s.excl_bits_off[e] : 0;
  const int64_t nw = (bits_words > 0 && s.excl_bits) ? s.excl_bits_off[e + 1] - w0 : 0;
  const bool have_bits = nw > 0 && nw <= bits_words;
  if (have_bits)
    for (int64_t i = threadIdx.x; i < nw; i += IPC_TPB) bits[i] = s.excl_bits[w0 + i];
  for (int k = threadIdx.x; k < IPC_CELLS + 2; k += IPC_TPB) cstart[k] = 0;
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o));
    mny = fmin(mny, __shfl_xor(mny, o));
    mxx = fmax(mxx, __shfl_xor(mxx, o));
    mxy = fmax(mxy, __shfl_xor(mxy, o));
  }
  const int w = threadIdx.x / WAVE;
  if (threadIdx.x % WAVE == 0) {
    red[4 * w] = mnx;
    red[4 * w + 1] = mny;
    red[4 * w + 2] = mxx;
    red[4 * w + 3] = mxy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < IPC_TPB / WAVE; ++k) {
      mnx = fmin(mnx, red[4 * k]);
      mny = fmin(mny, red[4 * k + 1]);
      mxx = fmax(mxx, red[4 * k + 2]);
      mxy = fmax(mxy, red[4 * k + 3]);
    }
    const double ex = fmax(mxx - mnx, 0.0), ey = fmax(mxy - mny, 0.0);
    double h = c.reach;
    while ((floor(ex / h) + 1.0) * (floor(ey / h) + 1.0) > double(IPC_CELLS)) h *= 1.4142135623730951;
    lo[0] = mnx;
    lo[1] = mny;
    cell = h;
    ncx = int(floor(ex / h)) + 1;
    ncy = int(floor(ey / h)) + 1;
  }
  __syncthreads();
  const double h = cell, x0 = lo[0], y0 = lo[1];
  const int nx = ncx, ny = ncy;
  // cell coordinates clamped in floating point first (a far or non-finite position lands in an
  // edge cell instead of overflowing the conversion)
  const auto cx_of = [&](double v) { return int(fmin(fmax(floor((v - x0) / h), 0.0), double(nx - 1))); };
  const auto cy_of = [&](double v) { return int(fmin(fmax(floor((v - y0) / h), 0.0), double(ny - 1))); };
  // counting sort of the boundary nodes by cell: counts into cstart[cell + 1], an inclusive
  // block scan, then each node counts its cell's end down to its slot, which leaves the start of
  // cell k in cstart[k + 1] (and n_be in cstart[ncell + 1]): cell k is [cstart[k+1], cstart[k+2]).
  const int ncell = nx * ny;
  for (int j = threadIdx.x; j < n_be; j += IPC_TPB)
    atomicAdd(cstart + cy_of(bp[2 * j + 1]) * nx + cx_of(bp[2 * j]) + 1, 1);
  __syncthreads();
  {
    __shared__ int wsum[IPC_TPB / WAVE];
    const int per = (ncell + IPC_TPB - 1) / IPC_TPB;
    const int k0 = 1 + threadIdx.x * per, k1 = min(k0 + per, ncell + 1);
    int sum = 0;
    for (int k = k0; k < k1; ++k) sum += cstart[k];
    const int lane = threadIdx.x % WAVE;
    int incl = sum;
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    int run = incl - sum;
    for (int q = 0; q < w; ++q) run += wsum[q];
    for (int k = k0; k < k1; ++k) {
      run += cstart[k];
      cstart[k] = run;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n_be; j += IPC_TPB) {
    const int cc = cy_of(bp[2 * j + 1]) * nx + cx_of(bp[2 * j]);
    citem[atomicAdd(cstart + cc + 1, -1) - 1] = j;
  }
  if (threadIdx.x == 0) cstart[ncell + 1] = n_be;
  __syncthreads();
  const BView bv{bp, slips ? bxn : nullptr, dirs ? bpd : nullptr, have_bits ? bits : nullptr};
  const double t_c = (mode == 1 || mode == 3) ? tc[e] : 0.0;
  // the gradient and energy passes (modes 0, 2) add terms only for pairs closer than dhat (a
  // pair's end distance is never below its segment distance), so they keep those; the
  // clearance and step-limit passes need every pair within the reach
  const double r = (mode == 1 || mode == 3 || mode == 4) ? c.reach : c.dhat;
  const int64_t eoff = s.excl_bits_off[e];
  const int ke0 = s.e_start[e], ne = s.e_start[e + 1] - ke0;
  double amin = BIG, cmin = BIG, acc = 0.0;     // this lane's, over all its pairs
  // Phase 1: this workgroup's edges (every ipc_parts(mode)-th block of IPC_TPB / IPC_TEAM), a
  // team of IPC_TEAM lanes per edge splitting the candidates of the cells the edge's box,
  // widened by r, overlaps; the pairs that pass the filter go to the LDS pair list (edge and
  // node packed in 16 bits each). When the list is full a lane works its pair at once.
  const int lane = threadIdx.x % IPC_TEAM, team = threadIdx.x / IPC_TEAM;
  __shared__ int npairs;
  if (threadIdx.x == 0) npairs = 0;
  __syncthreads();
  for (int base = blockIdx.y * (IPC_TPB / IPC_TEAM); base < ne; base += ipc_parts(mode) * (IPC_TPB / IPC_TEAM)) {
    const int kk = base + team;
    if (kk >= ne) continue;
    const EdgeData ed = edge_data(s, ke0 + kk, mode, bv, pdir);
    const int ax = cx_of(fmin(ed.xa.x, ed.xb.x) - r), bx = cx_of(fmax(ed.xa.x, ed.xb.x) + r);
    const int ay = cy_of(fmin(ed.xa.y, ed.xb.y) - r), by = cy_of(fmax(ed.xa.y, ed.xb.y) + r);
    int seen = 0;    // candidates in the cells before this one
    for (int gy = ay; gy <= by; ++gy)
      for (int gx = ax; gx <= bx; ++gx) {
        const int cc = gy * nx + gx;
        const int q0 = cstart[cc + 1], q1 = cstart[cc + 2];
        for (int q = q0 + ((lane - seen) % IPC_TEAM + IPC_TEAM) % IPC_TEAM; q < q1; q += IPC_TEAM) {
          const int jl = citem[q];
          const V2<double> xp = {bp[2 * jl], bp[2 * jl + 1]};
          if (!pair_in(s, ed, jl, xp, r, n_be, eoff, bv)) continue;
          const int slot = kk < 65536 ? atomicAdd(&npairs, 1) : plist_cap;
          if (slot < plist_cap) plist[slot] = (uint32_t(kk) << 16) | uint32_t(jl);
          else pair_body(s, ke0 + kk, mode, c, ed, jl, -1, xp, pdir, grad, diag, freev, t_c, s.mu_node, bv,
                         amin, cmin, acc);
        }
        seen += q1 - q0;
      }
  }
  __syncthreads();
  // Phase 2: the listed pairs, spread evenly over every lane of the workgroup.
  const int nl = min(npairs, plist_cap);
  for (int i = threadIdx.x; i < nl; i += IPC_TPB) {
    const uint32_t pk = plist[i];
    const int kk = int(pk >> 16), jl = int(pk & 0xFFFFu);
    const EdgeData ed = edge_data(s, ke0 + kk, mode, bv, pdir);
    const V2<double> xp = {bp[2 * jl], bp[2 * jl + 1]};
    pair_body(s, ke0 + kk, mode, c, ed, jl, -1, xp, pdir, grad, diag, freev, t_c, s.mu_node, bv, amin, cmin, acc);
  }
  // the workgroup's minima and sum, then one atomic each per workgroup
  if (mode == 1 || mode == 2 || mode == 3) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      amin = fmin(amin, __shfl_xor(amin, o));
      cmin = fmin(cmin, __shfl_xor(cmin, o));
      acc += __shfl_xor(acc, o);
    }
    __syncthreads();
    if (threadIdx.x % WAVE == 0) {
      red[4 * w] = amin;
      red[4 * w + 1] = cmin;
      red[4 * w + 2] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < IPC_TPB / WAVE; ++k) {
        amin = fmin(amin, red[4 * k]);
        cmin = fmin(cmin, red[4 * k + 1]);
        acc += red[4 * k + 2];
      }
      if (mode != 2) {
        if (cmin < BIG) atomic_min_d(clr + e, cmin);
        if (amin < BIG) atomic_min_d(alpha + e, amin);
      }
      if (mode != 3 && acc != 0.0) atomicAdd((mode == 1 ? s.pHp : E) + e, acc);
    }
  }
}

// ------------------------------------------------------------------------- "node" contact
// One workgroup per scene. Its boundary positions (and, as the mode needs, the slips over the
// step and the direction) are staged in LDS when they fit. Each thread visits the partners of its
// boundary nodes by a scan of the scene's boundary list in order; the reference finds the same
// partners through a hash grid (implicit/kernels.py:772-869, implicit/body_contact.py:37-167).
template <bool STAGED>
__global__ __launch_bounds__(256) void k_node_env(int mode, NodeArgs a) {
  extern __shared__ double sh[];
  const int e = blockIdx.x;
  const int start = a.bnd_start[e], n_be = a.bnd_start[e + 1] - start;
  const bool fric = mode <= 2 && a.fric;
  const bool dir = mode == 1 || mode == 3;
  double *sx = sh, *su = sh + 2 * a.max_nb, *sp = sh + 4 * a.max_nb;
  if (STAGED) {
    for (int j = threadIdx.x; j < n_be; j += blockDim.x) {
      const int n = a.bnd[start + j];
      const V2<double> x = ld2(a.x, n);
      sx[2 * j] = x.x;
      sx[2 * j + 1] = x.y;
      if (fric) {
        const V2<double> u = x - ld2(a.xn, n);
        su[2 * j] = u.x;
        su[2 * j + 1] = u.y;
      }
      if (dir) {
        const V2<double> q = ld2(a.pdir, n);
        sp[2 * j] = q.x;
        sp[2 * j + 1] = q.y;
      }
    }
    __syncthreads();
  }
  auto X = [&](int j) -> V2<double> {
    if (STAGED) return {sx[2 * j], sx[2 * j + 1]};
    return ld2(a.x, a.bnd[start + j]);
  };
  auto U = [&](int j) -> V2<double> {
    if (STAGED) return {su[2 * j], su[2 * j + 1]};
    const int n = a.bnd[start + j];
    return ld2(a.x, n) - ld2(a.xn, n);
  };
  auto P = [&](int j) -> V2<double> {
    if (STAGED) return {sp[2 * j], sp[2 * j + 1]};
    return ld2(a.pdir, a.bnd[start + j]);
  };
  const uint32_t *xw = a.excl_bits + a.excl_bits_off[e];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n_be; i += blockDim.x) {
    const int n = a.bnd[start + i];
    const V2<double> xa = X(i);
    const V2<double> ua = fric ? U(i) : V2<double>{0.0, 0.0};
    const V2<double> pa = dir ? P(i) : V2<double>{0.0, 0.0};
    V2<double> g = {0.0, 0.0}, gf = {0.0, 0.0};
    double hx = 0.0, hy = 0.0, fx = 0.0, fy = 0.0, val = 0.0, valf = 0.0;
    const int64_t row = (int64_t)i * n_be;
    for (int jl = 0; jl < n_be; ++jl) {
      if (jl == i || bit_of(xw, row + jl)) continue;
      NodeVisit v;
      if (!node_visit(xa, X(jl), a.r, v)) continue;
      if (mode == 0) node_grad_diag(v, a.k, a.clamp, g, hx, hy);
      else if (mode == 1) val += node_pHp(v, a.k, a.clamp, pa - P(jl));
      else if (mode == 2) val += node_energy(v, a.k);
      else g = g + node_hvp(v, a.k, pa - P(jl));
      if (fric) {
        const double mu = node_pair_mu(a.mu, a.per_body, a.mu_node, n, a.bnd[start + jl]);
        const V2<double> du = ua - U(jl);
        if (mode == 0) node_fric_grad_diag(v, a.k, mu, a.epsv, a.inv_dt, du, gf, fx, fy);
        else if (mode == 1) valf += node_fric_pHp(v, a.k, mu, a.epsv, a.inv_dt, du, pa - P(jl));
        else valf += node_fric_energy(v, a.k, mu, a.epsv, a.inv_dt, du);
      }
    }
    if (mode == 0) {
      atomic_add2(a.g, n, a.h2 * g);
      atomic_add2(a.dg, n, V2<double>{a.h2 * hx, a.h2 * hy});
      if (fric) {
        const double f = a.freev ? (a.freev[n] ? 1.0 : 0.0) : 1.0;
        atomic_add2(a.g, n, f * (a.h2 * gf));
        atomic_add2(a.dg, n, V2<double>{f * (a.h2 * fx), f * (a.h2 * fy)});
      }
    } else if (mode == 3) {
      atomic_add2(a.g, n, a.h2 * g);
    } else {
      acc += a.h2 * val + a.h2 * valf;
    }
  }
  if (mode == 1 || mode == 2) {
    const double w = wave_sum(acc);
    if (threadIdx.x % WAVE == 0 && w != 0.0) atomicAdd(a.acc + e, w);
  }
}

int launch_node_env(int mode, const NodeArgs &a, hipStream_t st) {
  if (a.B <= 0) return 0;
  // staged layout: positions, slips at 2 max_nb, direction at 4 max_nb (max_nb 0: unknown)
  const size_t lds = sizeof(double) * 6 * (size_t)a.max_nb;
  if (a.max_nb > 0 && lds <= 48 * 1024) {
    hipLaunchKernelGGL(k_node_env<true>, dim3(a.B), dim3(256), lds, st, mode, a);
  } else {
    hipLaunchKernelGGL(k_node_env<false>, dim3(a.B), dim3(256), 0, st, mode, a);
  }
  D2F_HIP(hipGetLastError());
  return 0;
}

// alpha <= 0.45 reach / p_inf (reference implicit/ipc_contact.py:300-305)
__global__ void k_far_cap(int B, const double *p_inf, double reach, double *alpha) {
  GRID_LOOP(e, B) {
    if (e < B && p_inf[e] > 0.0) alpha[e] = fmin(alpha[e], 0.45 * reach / p_inf[e]);
  }
}

// Kinematic sweep (reference implicit/ipc_contact.py:308-340)
__global__ void k_bc_motion(S s) {
  GRID_LOOP(b, s.n_bc) {
    const bool v = b < s.n_bc && !skip_node(s, s.bc_idx[b]);
    double l = 0.0;
    int e = 0;
    if (v) {
      const int i = s.bc_idx[b];
      const V2<double> d = ld2(s.bc_target, b) - ld2(s.x, i);
      st2(s.bc_q, i, d);
      l = len(d);
      e = s.node2env[i];
    }
    env_max(s.q_inf, e, l, v);
  }
}
__global__ void k_bc_advance(S s) {
  GRID_LOOP(b, s.n_bc) {
    if (b < s.n_bc && !skip_node(s, s.bc_idx[b])) {
      const int i = s.bc_idx[b];
      const double t = s.sweep_beta[s.node2env[i]];
      const V2<double> tg = ld2(s.bc_target, b);
      const V2<double> xi = t >= 1.0 ? tg : ld2(s.x, i) + t * (tg - ld2(s.x, i));
      st2(s.x, i, xi);
      st2(s.xhat, i, xi);
    }
  }
}
__global__ void k_bc_wake(S s) {
  GRID_LOOP(e, s.B) {
    if (e < s.B && s.q_inf[e] > 0.0 && !skip_env(s, e)) s.active[e] = 1;
  }
}
__global__ void k_fill(double *a, int n, double v) {
  GRID_LOOP(i, n) {
    if (i < n) a[i] = v;
  }
}
// Start of a contact pass that caps a step (modes 1 and 3): reset the cap to 1 when asked (the
// sweep), snapshot it into t_c (every pair's ACCD stops at the cap the pass started from) and
// clear the pass's minimum distance.
__global__ void k_ipc_begin(int B, double *tc, double *alpha, double *clr, int reset) {
  GRID_LOOP(e, B) {
    if (e < B) {
      if (reset) alpha[e] = 1.0;
      tc[e] = alpha[e];
      clr[e] = BIG;
    }
  }
}
// Per-iteration reset of the affine groups' reductions: gq_prev = gq, gq = Hx = Hy = 0.
__global__ void k_affine_reset(S s) {
  GRID_LOOP(j, 24 * s.n_ag) {
    if (j < 24 * s.n_ag) {
      const int g = j / 24, r = j % 24;
      if (r < 6) {
        s.ag_gq_prev[6 * g + r] = s.ag_gq[6 * g + r];
        s.ag_gq[6 * g + r] = 0.0;
      } else if (r < 15) {
        s.ag_Hx[9 * g + r - 6] = 0.0;
      } else {
        s.ag_Hy[9 * g + r - 15] = 0.0;
      }
    }
  }
}
// Per-iteration reset of the rigid-mode groups' reductions.
__global__ void k_rigid_reset(S s) {
  GRID_LOOP(j, 18 * s.n_rg) {
    if (j < 18 * s.n_rg) {
      const int g = j / 18, r = j % 18;
      if (r < 3) s.rg_acc[3 * g + r] = 0.0;
      else if (r < 6) s.rg_gq[3 * g + r - 3] = 0.0;
      else if (r < 9) s.rg_gq_prev[3 * g + r - 6] = 0.0;
      else s.rg_H[9 * g + r - 9] = 0.0;
    }
  }
}

// ------------------------------------------------------------------------- pressure, tendons
// reference implicit/actuation.py:89-123
__global__ void k_pressure(S s, int mode, double h2, double *E) {
  GRID_LOOP(k, s.n_pe) {
    if (k < s.n_pe) {
      const int a = s.pe_a[k], b = s.pe_b[k], ch = s.pe_ch[k];
      const double pr = s.ch_p[ch];
      if (mode == 0) {
        const V2<double> t = ld2(s.x, b) - ld2(s.x, a);
        const V2<double> g = (-0.5 * h2 * pr) * V2<double>{t.y, -t.x};
        atomic_add2(s.grad, a, g);
        atomic_add2(s.grad, b, g);
      } else if (mode == 1) {
        const V2<double> pa = ld2(s.p, a), pb = ld2(s.p, b);
        atomicAdd(s.pHp + s.ch2env[ch], -h2 * pr * cross(pa, pb));
      } else {
        const V2<double> ref = ld2(s.ch_ref, ch);
        const V2<double> ra = ld2(s.x, a) - ref, rb = ld2(s.x, b) - ref;
        atomicAdd(E + s.ch2env[ch], -h2 * pr * 0.5 * cross(ra, rb));
      }
    }
  }
}

// reference implicit/actuation.py:395-446: mode 0 gradient and diagonal at both ends, 1 p^T H p
// per environment, 2 energy per environment
__global__ void k_springs(S s, int mode, double h2, double *E) {
  GRID_LOOP(j, s.n_sp) {
    if (j < s.n_sp) {
      const int a = s.sp_a[j], b = s.sp_b[j];
      const double k = s.sp_k[j];
      const V2<double> xb = b >= 0 ? ld2(s.x, b) : ld2(s.sp_anchor, j);
      const SpringGeo q = spring_geo(ld2(s.x, a), xb, k, s.sp_L0[j]);
      if (mode == 0) {
        const V2<double> g = (h2 * q.tension) * q.n;
        const V2<double> hd = spring_diag(q, k, h2);
        atomic_add2(s.grad, a, g);
        atomic_add2(s.diag, a, hd);
        if (b >= 0) {
          atomic_add2(s.grad, b, -1.0 * g);
          atomic_add2(s.diag, b, hd);
        }
      } else if (mode == 1) {
        V2<double> dp = ld2(s.p, a);
        if (b >= 0) dp = dp - ld2(s.p, b);
        atomicAdd(s.pHp + s.sp2env[j], spring_pHp(q, k, h2, dp));
      } else {
        const double ext = q.len - s.sp_L0[j];
        atomicAdd(E + s.sp2env[j], 0.5 * h2 * k * ext * ext);
      }
    }
  }
}

// ------------------------------------------------------------------------- active strain
// Bm = Bm0 A^-1, area = area0 det A per element of each region, in both layouts (reference
// implicit/actuation.py:328-337)
__global__ void k_apply_activation(const int32_t *elem, const int32_t *ereg, int n_el, const double *Bm0,
                                   const double *area0, const double *Ainv, const double *detA,
                                   d2f_elem_consts o) {
  GRID_LOOP(j, n_el) {
    if (j < n_el) {
      const int t = elem[j], r = ereg[j];
      const double *B0 = Bm0 + 4 * t, *Ai = Ainv + 4 * r;
      const double B[4] = {B0[0] * Ai[0] + B0[1] * Ai[2], B0[0] * Ai[1] + B0[1] * Ai[3],
                           B0[2] * Ai[0] + B0[3] * Ai[2], B0[2] * Ai[1] + B0[3] * Ai[3]};
      const double A = area0[t] * detA[r];
      for (int c = 0; c < 4; ++c) o.Bm[4 * t + c] = B[c];
      o.area[t] = A;
      if (o.sc_pos) {
        const int q = o.sc_pos[t];
        for (int c = 0; c < 4; ++c) {
          o.sc_Bm[4 * q + c] = B[c];
          o.sc_Bm32[4 * q + c] = float(B[c]);
        }
        const double Amu = A * o.mu[t], Alam = A * o.lam[t];
        o.sc_Amu[q] = Amu;
        o.sc_Alam[q] = Alam;
        o.sc_Amu32[q] = float(Amu);
        o.sc_Alam32[q] = float(Alam);
      }
    }
  }
}

__device__ __forceinline__ V2<double> seg_dir(V2<double> a, V2<double> b, double &il) {
  const V2<double> t = b - a;
  const double l = fmax(len(t), 1.0e-12);
  il = 1.0 / l;
  return {t.x / l, t.y / l};
}
// reference implicit/actuation.py:170-195
__global__ void k_tendon_length(S s) {
  GRID_LOOP(j, s.n_ey) {
    if (j < s.n_ey && s.ey_nxt[j] >= 0)
      atomicAdd(s.td_L + s.ey_tid[j], len(ld2(s.x, s.ey_node[s.ey_nxt[j]]) - ld2(s.x, s.ey_node[j])));
  }
}
__global__ void k_tendon_state(S s) {
  GRID_LOOP(k, s.n_td) {
    if (k < s.n_td) {
      if (s.td_stiff[k] > 0.0) {
        const double ext = s.td_L[k] - s.td_L0[k];
        s.td_tau[k] = ext > 0.0 ? s.td_stiff[k] * ext : 0.0;
        s.td_keff[k] = ext > 0.0 ? s.td_stiff[k] : 0.0;
      } else {
        s.td_tau[k] = fmax(s.td_T[k], 0.0);
        s.td_keff[k] = 0.0;
      }
    }
  }
}
// reference implicit/actuation.py:205-226
__global__ void k_tendon_grad(S s, double h2) {
  GRID_LOOP(j, s.n_ey) {
    if (j < s.n_ey) {
      const int n = s.ey_node[j], k = s.ey_tid[j];
      V2<double> gL = {0.0, 0.0}, geo = {0.0, 0.0};
      const V2<double> xj = ld2(s.x, n);
      double il;
      if (s.ey_prv[j] >= 0) {
        const V2<double> d = seg_dir(ld2(s.x, s.ey_node[s.ey_prv[j]]), xj, il);
        gL = gL + d;
        geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
      }
      if (s.ey_nxt[j] >= 0) {
        const V2<double> d = seg_dir(xj, ld2(s.x, s.ey_node[s.ey_nxt[j]]), il);
        gL = gL - d;
        geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
      }
      const double tau = s.td_tau[k], ke = s.td_keff[k];
      atomic_add2(s.grad, n, (h2 * tau) * gL);
      atomic_add2(s.diag, n, V2<double>{h2 * (tau * geo.x + ke * gL.x * gL.x),
                                        h2 * (tau * geo.y + ke * gL.y * gL.y)});
    }
  }
}
// reference implicit/actuation.py:229-255
__global__ void k_tendon_pHp(S s, double h2) {
  GRID_LOOP(j, s.n_ey) {
    if (j < s.n_ey && s.ey_nxt[j] >= 0) {
      const int k = s.ey_tid[j];
      const int na = s.ey_node[j], nb = s.ey_node[s.ey_nxt[j]];
      double il;
      const V2<double> d = seg_dir(ld2(s.x, na), ld2(s.x, nb), il);
      const V2<double> dp = ld2(s.p, nb) - ld2(s.p, na);
      const double along = dot(dp, d);
      atomicAdd(s.td_dL + k, along);
      atomicAdd(s.pHp + s.td2env[k], h2 * s.td_tau[k] * (dot(dp, dp) - along * along) * il);
    }
  }
}
__global__ void k_tendon_taut(S s, double h2) {
  GRID_LOOP(k, s.n_td) {
    if (k < s.n_td) atomicAdd(s.pHp + s.td2env[k], h2 * s.td_keff[k] * s.td_dL[k] * s.td_dL[k]);
  }
}
// reference implicit/actuation.py:258-268
__global__ void k_tendon_energy(S s, double h2, double *E) {
  GRID_LOOP(k, s.n_td) {
    if (k < s.n_td) {
      if (s.td_stiff[k] > 0.0) {
        const double ext = s.td_L[k] - s.td_L0[k];
        if (ext > 0.0) atomicAdd(E + s.td2env[k], 0.5 * h2 * s.td_stiff[k] * ext * ext);
      } else {
        atomicAdd(E + s.td2env[k], h2 * fmax(s.td_T[k], 0.0) * s.td_L[k]);
      }
    }
  }
}

// ------------------------------------------------------------------------- affine bodies
// reference implicit/affine.py:61-269
// The slaved nodes are grouped by body, so a wave's lanes almost always share a group: each
// sum is reduced across the wave first and lands as one atomic per wave (env_add).
__global__ void k_affine_gather(S s) {
  GRID_LOOP(k, s.n_an) {
    const bool ok = k < s.n_an;
    const int i = ok ? s.an_node[k] : 0, b = ok ? s.an_grp[k] : -1;
    double w[3] = {0.0, 0.0, 0.0};
    V2<double> g = {0.0, 0.0}, d = {0.0, 0.0};
    if (ok) {
      w[0] = s.an_w[3 * k];
      w[1] = s.an_w[3 * k + 1];
      w[2] = s.an_w[3 * k + 2];
      g = ld2(s.grad, i);
      d = ld2(s.diag, i);
    }
    for (int r = 0; r < 3; ++r) {
      env_add(s.ag_gq + r, 6 * b, g.x * w[r], ok);
      env_add(s.ag_gq + 3 + r, 6 * b, g.y * w[r], ok);
      for (int c = 0; c < 3; ++c) {
        env_add(s.ag_Hx + 3 * r + c, 9 * b, d.x * w[r] * w[c], ok);
        env_add(s.ag_Hy + 3 * r + c, 9 * b, d.y * w[r] * w[c], ok);
      }
    }
    if (ok) st2(s.grad, i, V2<double>{0.0, 0.0});
  }
}

__global__ void k_affine_reduce(S s, double h2) {
  GRID_LOOP(b, s.n_ag) {
    if (b < s.n_ag) {
      const double *q = s.ag_q + 6 * b;
      const double Mu = s.ag_Mu[b], La = s.ag_La[b];
      const NH<double> h = nh_eval(M2<double>{q[1], q[2], q[4], q[5]}, Mu, La);
      const double P[4] = {Mu * h.F.a + h.c * h.G.a, Mu * h.F.b + h.c * h.G.b,
                           Mu * h.F.c + h.c * h.G.c, Mu * h.F.d + h.c * h.G.d};
      double g[6], y[6], z[6], zy[6];
      const double add[6] = {0.0, P[0], P[1], 0.0, P[2], P[3]};
      for (int r = 0; r < 6; ++r) g[r] = s.ag_gq[6 * b + r] + h2 * add[r];
      M66 H, L;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) H.m[r][c] = 0.0;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          H.m[r][c] = s.ag_Hx[9 * b + 3 * r + c];
          H.m[r + 3][c + 3] = s.ag_Hy[9 * b + 3 * r + c];
        }
      affine_block(H, q[1], q[2], q[4], q[5], Mu, La, h2);
      cholesky6(H, L);
      for (int r = 0; r < 6; ++r) y[r] = g[r] - s.ag_gq_prev[6 * b + r];
      cholesky6_solve(L, g, z);
      cholesky6_solve(L, y, zy);
      double gz = 0, gp = 0, gzy = 0, yp = 0, yzy = 0;
      for (int r = 0; r < 6; ++r) {
        const double pp = s.ag_pq[6 * b + r];
        gz += g[r] * z[r];
        gp += g[r] * pp;
        gzy += g[r] * zy[r];
        yp += y[r] * pp;
        yzy += y[r] * zy[r];
      }
      const int e = s.ag2env[b];
      atomicAdd(s.gPg + e, gz);
      atomicAdd(s.g_p + e, gp);
      atomicAdd(s.g_Py + e, gzy);
      atomicAdd(s.y_p + e, yp);
      atomicAdd(s.y_Py + e, yzy);
      for (int r = 0; r < 6; ++r) {
        s.ag_gq[6 * b + r] = g[r];
        s.ag_zq[6 * b + r] = z[r];
      }
    }
  }
}

// which = 0: pq = -zq + beta pq; 1: q += alpha pq (active); 2: q = q0 + alpha pq (active)
__global__ void k_affine_update(S s, int which) {
  GRID_LOOP(b, s.n_ag) {
    if (b < s.n_ag) {
      const int e = s.ag2env[b];
      for (int r = 0; r < 6; ++r) {
        const int k = 6 * b + r;
        if (which == 0) s.ag_pq[k] = -s.ag_zq[k] + s.beta[e] * s.ag_pq[k];
        else if (which == 1) { if (s.active[e]) s.ag_q[k] = s.ag_q[k] + s.alpha[e] * s.ag_pq[k]; }
        else { if (s.active[e]) s.ag_q[k] = s.ag_q0[k] + s.alpha[e] * s.ag_pq[k]; }
      }
    }
  }
}

__global__ void k_affine_expand(S s, const double *v6, double *out) {
  GRID_LOOP(k, s.n_an) {
    if (k < s.n_an && !(out == s.x && skip_node(s, s.an_node[k]))) {
      const double *w = s.an_w + 3 * k;
      const double *v = v6 + 6 * s.an_grp[k];
      st2(out, s.an_node[k], V2<double>{w[0] * v[0] + w[1] * v[1] + w[2] * v[2],
                                        w[0] * v[3] + w[1] * v[4] + w[2] * v[5]});
    }
  }
}

__global__ void k_affine_pHp_energy(S s, int mode, double h2, double *E) {
  GRID_LOOP(b, s.n_ag) {
    if (b < s.n_ag) {
      const double *q = s.ag_q + 6 * b;
      const double Mu = s.ag_Mu[b], La = s.ag_La[b];
      const NH<double> h = nh_eval(M2<double>{q[1], q[2], q[4], q[5]}, Mu, La);
      if (mode == 1) {
        const double *pq = s.ag_pq + 6 * b;
        const double v = nh_quad(h, M2<double>{pq[1], pq[2], pq[4], pq[5]}, Mu, La);
        atomicAdd(s.pHp + s.ag2env[b], h2 * fmax(v, 0.0));
      } else {
        atomicAdd(E + s.ag2env[b], h2 * nh_psi(h, Mu, La));
      }
    }
  }
}

// ------------------------------------------------------------------------- rigid modes
// reference implicit/affine.py:405-486
__global__ void k_rigid_centroid(S s) {
  GRID_LOOP(k, s.n_rn) {
    const bool ok = k < s.n_rn;
    const int i = ok ? s.rn_node[k] : 0, b = ok ? s.rn_grp[k] : -1;
    const double m = ok ? s.mass[i] : 0.0;
    const V2<double> x = ok ? ld2(s.x, i) : V2<double>{0.0, 0.0};
    env_add(s.rg_acc, 3 * b, m * x.x, ok);
    env_add(s.rg_acc + 1, 3 * b, m * x.y, ok);
    env_add(s.rg_acc + 2, 3 * b, m, ok);
  }
}
__device__ __forceinline__ V2<double> rigid_arm(const S &s, int i, int b) {
  const double *a = s.rg_acc + 3 * b;
  const V2<double> x = ld2(s.x, i);
  return {x.x - a[0] / a[2], x.y - a[1] / a[2]};
}
__global__ void k_rigid_gather(S s) {
  GRID_LOOP(k, s.n_rn) {
    const bool ok = k < s.n_rn;
    const int i = ok ? s.rn_node[k] : 0, b = ok ? s.rn_grp[k] : -1;
    V2<double> rr = {0.0, 0.0}, g = {0.0, 0.0}, gp = {0.0, 0.0};
    double dx = 0.0, dy = 0.0;
    if (ok) {
      rr = rigid_arm(s, i, b);
      g = ld2(s.grad, i);
      gp = ld2(s.grad_prev, i);
      const double m = s.mass[i];
      const V2<double> d = ld2(s.diag, i) - ld2(s.diag_el, i);
      dx = fmax(d.x, 0.0) + m;
      dy = fmax(d.y, 0.0) + m;
    }
    env_add(s.rg_gq, 3 * b, g.x, ok);
    env_add(s.rg_gq + 1, 3 * b, g.y, ok);
    env_add(s.rg_gq + 2, 3 * b, rr.x * g.y - rr.y * g.x, ok);
    env_add(s.rg_gq_prev, 3 * b, gp.x, ok);
    env_add(s.rg_gq_prev + 1, 3 * b, gp.y, ok);
    env_add(s.rg_gq_prev + 2, 3 * b, rr.x * gp.y - rr.y * gp.x, ok);
    env_add(s.rg_H, 9 * b, dx, ok);
    env_add(s.rg_H + 2, 9 * b, -dx * rr.y, ok);
    env_add(s.rg_H + 4, 9 * b, dy, ok);
    env_add(s.rg_H + 5, 9 * b, dy * rr.x, ok);
    env_add(s.rg_H + 6, 9 * b, -dx * rr.y, ok);
    env_add(s.rg_H + 7, 9 * b, dy * rr.x, ok);
    env_add(s.rg_H + 8, 9 * b, dx * rr.y * rr.y + dy * rr.x * rr.x, ok);
  }
}
__global__ void k_rigid_reduce(S s) {
  GRID_LOOP(b, s.n_rg) {
    if (b < s.n_rg) {
      double Hi[9];
      invert33(s.rg_H + 9 * b, Hi);
      const double *g = s.rg_gq + 3 * b, *g0 = s.rg_gq_prev + 3 * b;
      const double y[3] = {g[0] - g0[0], g[1] - g0[1], g[2] - g0[2]};
      double z[3], zy[3];
      for (int r = 0; r < 3; ++r) {
        z[r] = Hi[3 * r] * g[0] + Hi[3 * r + 1] * g[1] + Hi[3 * r + 2] * g[2];
        zy[r] = Hi[3 * r] * y[0] + Hi[3 * r + 1] * y[1] + Hi[3 * r + 2] * y[2];
      }
      const int e = s.rg2env[b];
      atomicAdd(s.gPg + e, g[0] * z[0] + g[1] * z[1] + g[2] * z[2]);
      atomicAdd(s.g_Py + e, g[0] * zy[0] + g[1] * zy[1] + g[2] * zy[2]);
      atomicAdd(s.y_Py + e, y[0] * zy[0] + y[1] * zy[1] + y[2] * zy[2]);
      for (int r = 0; r < 3; ++r) s.rg_zq[3 * b + r] = z[r];
    }
  }
}
__global__ void k_rigid_direction(S s) {
  GRID_LOOP(k, s.n_rn) {
    if (k < s.n_rn) {
      const int i = s.rn_node[k], b = s.rn_grp[k];
      const V2<double> rr = rigid_arm(s, i, b);
      const double *z = s.rg_zq + 3 * b;
      st2(s.p, i, ld2(s.p, i) - V2<double>{z[0] - z[2] * rr.y, z[1] + z[2] * rr.x});
    }
  }
}

// ------------------------------------------------------------------------- direction
// reference implicit/kernels.py:161-163, 168-192
__global__ void k_mask_dk(S s) {
  GRID_LOOP(i, s.N) {
    const bool v = i < s.N;
    double a = 0, b = 0, c = 0, d = 0, f = 0;
    int e = 0;
    if (v) {
      e = s.node2env[i];
      const double fr = fr_of(s, i);
      V2<double> g = ld2(s.grad, i);
      g = {g.x * fr, g.y * fr};
      st2(s.grad, i, g);
      const V2<double> dh = ld2(s.diag, i);
      const double Px = fr / fmax(dh.x, 1.0e-12), Py = fr / fmax(dh.y, 1.0e-12);
      const V2<double> pp = ld2(s.p, i);
      const V2<double> yy = g - ld2(s.grad_prev, i);
      const double Pg0 = Px * g.x, Pg1 = Py * g.y, Py0 = Px * yy.x, Py1 = Py * yy.y;
      a = g.x * Pg0 + g.y * Pg1;
      b = g.x * pp.x + g.y * pp.y;
      c = g.x * Py0 + g.y * Py1;
      d = yy.x * pp.x + yy.y * pp.y;
      f = yy.x * Py0 + yy.y * Py1;
    }
    env_add(s.gPg, e, a, v);
    env_add(s.g_p, e, b, v);
    env_add(s.g_Py, e, c, v);
    env_add(s.y_p, e, d, v);
    env_add(s.y_Py, e, f, v);
  }
}

// reference implicit/kernels.py:196-217 and implicit/solver.py:77-93
__global__ void k_finalize_beta(S s, int it, double guard, int use_hold) {
  GRID_LOOP(e, s.B) {
    if (e < s.B) {
      double b = 0.0;
      if (it > 0) {
        const double ytp = s.y_p[e];
        if (fabs(ytp) > 1.0e-30) b = s.g_Py[e] / ytp - s.y_Py[e] * s.g_p[e] / (ytp * ytp);
      }
      double t = -s.gPg[e] + b * s.g_p[e];
      if (t >= 0.0) {
        b = 0.0;
        t = -s.gPg[e];
      }
      if (guard > 0.0 && it > 0) {
        const double hold = use_hold ? s.q_inf[e] : 0.0;
        if (hold <= 0.0 && s.y_p[e] < guard * s.alpha[e] * s.pHp[e]) {
          b = 0.0;
          t = -s.gPg[e];
        }
      }
      s.beta[e] = b;
      s.gTp[e] = t;
    }
  }
}

// reference implicit/kernels.py:221-230
__global__ void k_update_p(S s) {
  GRID_LOOP(i, s.N) {
    if (i < s.N) {
      const double fr = fr_of(s, i);
      const double bb = s.beta[s.node2env[i]];
      const V2<double> dh = ld2(s.diag, i), g = ld2(s.grad, i), pp = ld2(s.p, i);
      const double px = -fr / fmax(dh.x, 1.0e-12) * g.x + bb * pp.x;
      const double py = -fr / fmax(dh.y, 1.0e-12) * g.y + bb * pp.y;
      st2(s.p, i, V2<double>{px * fr, py * fr});
    }
  }
}

// ------------------------------------------------------------------------- curvature, caps
// reference implicit/kernels.py:235-255
__global__ void k_pHp_reset(S s, double alpha_max) {
  GRID_LOOP(e, s.B) {
    if (e < s.B) {
      s.pHp[e] = 0.0;
      s.alpha[e] = alpha_max;
      s.p_inf[e] = 0.0;
      s.clr_min[e] = BIG;
    }
  }
}
__global__ void k_pHp_mass(S s) {
  GRID_LOOP(i, s.N) {
    const bool v = i < s.N;
    double a = 0.0, l = 0.0;
    int e = 0;
    if (v) {
      e = s.node2env[i];
      const V2<double> pp = ld2(s.p, i);
      a = s.mass[i] * dot(pp, pp);
      l = len(pp);
    }
    env_add(s.pHp, e, a, v);
    env_max(s.p_inf, e, l, v);
  }
}

__global__ void k_elastic_pHp(const double *x, const double *p, const int32_t *tri, const double *Bm,
                              const double *area, const double *mu, const double *lam,
                              const int32_t *tri2env, int T, double h2, int clamp, double *pHp) {
  GRID_LOOP(t, T) {
    const bool v = t < T;
    double val = 0.0;
    int e = 0;
    if (v) {
      const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      const V2<double> p0 = ld2(x, i0), p1 = ld2(x, i1), p2 = ld2(x, i2);
      const M2<double> B = load_B<double>(Bm, t);
      const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
      const NH<double> h = nh_eval(mat_mul(Ds, B), mu[t], lam[t]);
      const M2<double> dF = dF_of(ld2(p, i0), ld2(p, i1), ld2(p, i2), B);
      val = area[t] * nh_quad(h, dF, mu[t], lam[t]);
      if (clamp) val = fmax(val, 0.0);
      val *= h2;
      e = tri2env[t];
    }
    env_add(pHp, e, val, v);
  }
}

__global__ void k_inversion_alpha(const double *x, const double *p, const int32_t *tri,
                                  const int32_t *tri2env, int T, double shrink, double *alpha) {
  GRID_LOOP(t, T) {
    bool v = t < T;
    double sv = BIG;
    int e = 0;
    if (v) {
      const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      const V2<double> x0 = ld2(x, i0), p0 = ld2(p, i0);
      sv = inversion_root(ld2(x, i1) - x0, ld2(x, i2) - x0, ld2(p, i1) - p0, ld2(p, i2) - p0, shrink);
      e = tri2env[t];
      v = sv < BIG;
    }
    env_min(alpha, e, sv, v);
  }
}

__global__ void k_elastic_energy(const double *x, const int32_t *tri, const double *Bm,
                                 const double *area, const double *mu, const double *lam,
                                 const int32_t *tri2env, int T, double h2, double *E) {
  GRID_LOOP(t, T) {
    const bool v = t < T;
    double val = 0.0;
    int e = 0;
    if (v) {
      const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      const V2<double> p0 = ld2(x, i0), p1 = ld2(x, i1), p2 = ld2(x, i2);
      const M2<double> B = load_B<double>(Bm, t);
      const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
      const NH<double> h = nh_eval(mat_mul(Ds, B), mu[t], lam[t]);
      val = h2 * area[t] * nh_psi(h, mu[t], lam[t]);
      e = tri2env[t];
    }
    env_add(E, e, val, v);
  }
}

__global__ void k_elastic_hvp(const double *x, const double *vv, const int32_t *tri, const double *Bm,
                              const double *area, const double *mu, const double *lam, int T,
                              double h2, double *y) {
  GRID_LOOP(t, T) {
    if (t < T) {
      const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
      const V2<double> p0 = ld2(x, i0), p1 = ld2(x, i1), p2 = ld2(x, i2);
      const M2<double> B = load_B<double>(Bm, t);
      const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
      const NH<double> h = nh_eval(mat_mul(Ds, B), mu[t], lam[t]);
      const M2<double> dF = dF_of(ld2(vv, i0), ld2(vv, i1), ld2(vv, i2), B);
      const M2<double> dP = nh_d2_dir(h, dF, mu[t], lam[t]);
      const double A = area[t];
      const V2<double> h1 = {A * (dP.a * B.a + dP.b * B.b), A * (dP.c * B.a + dP.d * B.b)};
      const V2<double> hh2 = {A * (dP.a * B.c + dP.b * B.d), A * (dP.c * B.c + dP.d * B.d)};
      atomic_add2(y, i0, h2 * V2<double>{-(h1.x + hh2.x), -(h1.y + hh2.y)});
      atomic_add2(y, i1, h2 * h1);
      atomic_add2(y, i2, h2 * hh2);
    }
  }
}

// Collider curvature + clearance, friction curvature (core/sdf.py:158-171, 237-255)
__global__ void k_colliders_pHp(S s, Pm p) {
  GRID_LOOP(b, s.n_b) {
    const bool v = b < s.n_b;
    double acc = 0.0, cl = BIG;
    int e = 0;
    if (v) {
      const int n = s.bnd[b];
      e = s.b2e[b];
      const V2<double> x = ld2(s.x, n), pp = ld2(s.p, n);
      for (int k = 0; k < s.n_col; ++k) {
        double dist;
        acc += collider_pHp(to_collider(s.col[k]), x, pp, p.dhat, p.kappa, p.h2, 1, dist);
        cl = fmin(cl, dist);
      }
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          const double kt = p.h2 * c.mu * f.f_n * p.inv_dt / fmax(f.vtm, p.epsv);
          const double pt = dot(pp, V2<double>{-f.n.y, f.n.x});
          acc += kt * pt * pt;
        }
      }
    }
    env_add(s.pHp, e, acc, v);
    env_min(s.clr_min, e, cl, v && s.n_col > 0);
  }
}

__global__ void k_colliders_ccd(S s, double eta) {
  GRID_LOOP(b, s.n_b) {
    bool v = b < s.n_b;
    double a = BIG;
    int e = 0;
    if (v) {
      const int n = s.bnd[b];
      e = s.b2e[b];
      const V2<double> x = ld2(s.x, n), pp = ld2(s.p, n);
      for (int k = 0; k < s.n_col; ++k) a = fmin(a, collider_ccd(to_collider(s.col[k]), x, pp, eta));
      v = a < BIG;
    }
    env_min(s.alpha, e, a, v);
  }
}

__global__ void k_colliders_energy(S s, Pm p, double *E) {
  GRID_LOOP(b, s.n_b) {
    const bool v = b < s.n_b;
    double acc = 0.0;
    int e = 0;
    if (v) {
      const int n = s.bnd[b];
      e = s.b2e[b];
      const V2<double> x = ld2(s.x, n);
      for (int k = 0; k < s.n_col; ++k) acc += collider_energy(to_collider(s.col[k]), x, p.dhat, p.kappa, p.h2);
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          const double ut = f.vtm / p.inv_dt, a = p.epsv / p.inv_dt;
          const double f0 = ut < a ? ut * ut / (2.0 * a) : ut - 0.5 * a;
          acc += p.h2 * c.mu * f.f_n * f0;
        }
      }
    }
    env_add(E, e, acc, v);
  }
}

// inertia + external force energy (implicit/kernels.py:446-451, implicit/actuation.py:44-50)
__global__ void k_node_energy(S s, double h2, double *E) {
  GRID_LOOP(i, s.N) {
    const bool v = i < s.N;
    double val = 0.0;
    int e = 0;
    if (v) {
      e = s.node2env[i];
      const V2<double> d = ld2(s.x, i) - ld2(s.xhat, i);
      val = 0.5 * s.mass[i] * dot(d, d);
      if (s.fext) val += -h2 * dot(ld2(s.fext, i), ld2(s.x, i) - ld2(s.xn, i));
    }
    env_add(E, e, val, v);
  }
}

// reference implicit/kernels.py:375-423
__global__ void k_alpha_finalize(S s, Pm p, int it) {
  GRID_LOOP(e, s.B) {
    if (e < s.B && s.active[e] != 0) {
      const double ph = s.pHp[e];
      double a = s.alpha[e];
      if (ph > 0.0) a = fmin(a, -s.gTp[e] / ph);
      if (s.clr_min[e] < p.near_factor * p.dhat && s.p_inf[e] > 0.0) a = fmin(a, 0.5 * p.dhat / s.p_inf[e]);
      if (a <= 0.0) {
        s.active[e] = 0;
        s.alpha[e] = 0.0;
      } else {
        s.alpha[e] = a;
        if (p.tol_x > 0.0 && a * s.p_inf[e] < p.tol_x) {
          s.active[e] = 0;
          s.iters[e] += 1;
        } else {
          const double d = -a * s.gTp[e] - 0.5 * a * a * ph;
          s.dE[e] = d;
          if (it == 0) s.dE0[e] = fmax(d, 1.0e-300);
          else if (d < p.eps * s.dE0[e]) s.active[e] = 0;
          s.iters[e] += 1;
        }
      }
    }
  }
}

// x += alpha p (active), or x = x0 + alpha p for a line-search trial
// (implicit/kernels.py:427-434, implicit/solver.py:52-60)
__global__ void k_update_x(S s, const double *from) {
  GRID_LOOP(i, s.N) {
    if (i < s.N) {
      const int e = s.node2env[i];
      if (s.active[e]) st2(s.x, i, ld2(from, i) + s.alpha[e] * ld2(s.p, i));
    }
  }
}

// reference implicit/solver.py:64-73
__global__ void k_ls_halve(S s, double tol) {
  GRID_LOOP(e, s.B) {
    if (e < s.B && s.active[e] && s.Ener[e] > s.Eprev[e] + tol * (fabs(s.Eprev[e]) + 1.0))
      s.alpha[e] *= 0.5;
  }
}

// ------------------------------------------------------------------------- host driver
static inline int nblk(int n) { return n <= 0 ? 1 : (n + TPB - 1) / TPB; }

#define LAUNCH(kern, n, ...)                                                          \
  do {                                                                                \
    if ((n) > 0) hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(TPB), 0, st, __VA_ARGS__); \
  } while (0)

struct Plain {
  const S &s;
  const Pm &p;
  hipStream_t st;
  PairCtx pc;
  bool pending = false;
  std::vector<double> hbuf;
  std::vector<int> ibuf;

  Plain(const S &s_, const Pm &p_, hipStream_t st_) : s(s_), p(p_), st(st_) {
    pc = PairCtx{p.dhat, p.ipc_kappa, p.ipc_mu, p.epsv, p.inv_dt, p.h2, p.ipc_reach, p.ipc_gap, 0.0,
                 p.ipc_per_body, nullptr, 0};
  }

  int ipc_pass(int mode, const double *pdir, double *grad, double *diag, const uint8_t *freev,
               double *E, double *clr, double *alpha, double h2, double mu) {
    PairCtx c = pc;
    c.h2 = h2;
    c.mu = mu;
    c.floor_ = mode == 3 ? 0.25 * p.dhat : 0.0;
    // modes 1 and 3 read the t_c snapshot that k_ipc_begin took before the pass
    const size_t base = ipc_env_lds(mode, s.max_nb);
    // 64 KiB with the static LDS; edge and node indices packed in 16 bits in the pair list
    if (s.e_start && s.max_nb > 0 && s.max_nb < 65536 && base <= 52 * 1024) {
      // the largest exclusion bitset that still fits, staged by the scenes whose bitset it holds;
      // the rest of the 60 KiB holds the pair list (at least 1024 pairs)
      const bool reach_pass = mode == 1 || mode == 3 || mode == 4;   // many candidates: stage the bitset
      const int words = (s.excl_bits && reach_pass) ? (int)std::min<size_t>((56 * 1024 - base) / 4, (size_t)s.excl_bits_lds) : 0;
      const int cap = (int)((60 * 1024 - base - 4 * (size_t)words) / 4);
      const size_t lds = base + 4 * (size_t)words + 4 * (size_t)cap;
      const dim3 grid(s.B, ipc_parts(mode));
#define D2F_IPC_ENV(M)                                                                                  \
  hipLaunchKernelGGL(k_ipc_env<M>, grid, dim3(IPC_TPB), lds, st, s, c, pdir, grad, diag, freev, E, clr, \
                     alpha, s.tc, words, cap)
      if (mode == 0) D2F_IPC_ENV(0);
      else if (mode == 1) D2F_IPC_ENV(1);
      else if (mode == 2) D2F_IPC_ENV(2);
      else if (mode == 3) D2F_IPC_ENV(3);
      else D2F_IPC_ENV(4);
#undef D2F_IPC_ENV
      D2F_HIP(hipGetLastError());
    } else {
      LAUNCH(k_ipc, s.n_e, s, mode, c, pdir, grad, diag, freev, E, clr, alpha, s.tc);
    }
    return 0;
  }

  // "node" contact pass (reference implicit/solver.py:279-285, 370-375, 475-481)
  int node_pass(int mode, double h2, double *acc) {
    NodeArgs a{};
    a.x = s.x;
    a.xn = s.xn;
    a.pdir = s.p;
    a.bnd = s.bnd;
    a.bnd_start = s.bnd_start;
    a.B = s.B;
    a.max_nb = s.max_nb;
    a.excl_bits = s.excl_bits;
    a.excl_bits_off = s.excl_bits_off;
    a.freev = s.free_;
    a.mu_node = s.mu_node;
    a.k = p.node_k;
    a.r = p.node_r;
    a.mu = p.node_mu;
    a.epsv = p.epsv;
    a.inv_dt = p.inv_dt;
    a.h2 = h2;
    a.clamp = 1;
    a.fric = p.node_fric;
    a.per_body = p.ipc_per_body;
    a.g = s.grad;
    a.dg = s.diag;
    a.acc = acc;
    return launch_node_env(mode, a, st);
  }

  int ccd_sweep() {  // IPCContact.ccd along bc_q into sweep_beta (reference ipc_contact.py:432-439)
    LAUNCH(k_ipc_begin, s.B, s.B, s.tc, s.sweep_beta, s.clr_ipc, 1);
    int r = ipc_pass(3, s.bc_q, s.grad, s.diag, s.free_, s.Ener, s.clr_ipc, s.sweep_beta, p.h2, p.ipc_mu);
    if (r) return r;
    LAUNCH(k_far_cap, s.B, s.B, s.q_inf, p.ipc_reach, s.sweep_beta);
    return 0;
  }

  int sweep() {
    LAUNCH(k_fill, s.B, s.q_inf, s.B, 0.0);
    LAUNCH(k_bc_motion, s.n_bc, s);
    int r = ccd_sweep();
    if (r) return r;
    LAUNCH(k_bc_advance, s.n_bc, s);
    return 0;
  }

  int poll_pending(bool &out) {
    hbuf.resize(s.B);
    D2F_HIP(hipMemcpyAsync(hbuf.data(), s.sweep_beta, sizeof(double) * s.B, hipMemcpyDeviceToHost, st));
    D2F_HIP(hipStreamSynchronize(st));
    out = false;
    for (int e = 0; e < s.B; ++e)
      if (!(hbuf[e] >= 1.0)) out = true;
    return 0;
  }

  int grad_diag() {
    LAUNCH(k_grad_diag_mass, s.N, s, p.h2);
    {
      const int rc = launch_elastic_grad_diag(s.x, s.tri, s.Bm, s.area, s.mu, s.lam, s.T, p.h2, 1, &s.el,
                                              s.grad, s.diag, st);
      if (rc) return rc;
    }
    if (s.n_rg > 0)
      D2F_HIP(hipMemcpyAsync(s.diag_el, s.diag, sizeof(double) * 2 * s.N, hipMemcpyDeviceToDevice, st));
    if (s.n_col > 0) LAUNCH(k_colliders_grad_diag, s.n_b, s, p);
    if (p.node_on) {
      const int r = node_pass(0, p.h2, nullptr);
      if (r) return r;
    }
    // terms, in the reference's order: external force, affine (none), IPC, pressure, tendon
    if (s.fext) LAUNCH(k_extforce_grad, s.N, s, p.h2);
    if (p.ipc_on) {
      int r = ipc_pass(0, s.p, s.grad, s.diag, s.free_, s.Ener, s.clr_min, s.alpha, p.h2, p.ipc_mu);
      if (r) return r;
    }
    if (s.n_sp > 0) LAUNCH(k_springs, s.n_sp, s, 0, p.h2, s.Ener);
    if (s.n_pe > 0) LAUNCH(k_pressure, s.n_pe, s, 0, p.h2, s.Ener);
    if (s.n_td > 0) {
      LAUNCH(k_fill, s.n_td, s.td_L, s.n_td, 0.0);
      LAUNCH(k_tendon_length, s.n_ey, s);
      LAUNCH(k_tendon_state, s.n_td, s);
      LAUNCH(k_tendon_grad, s.n_ey, s, p.h2);
    }
    // modes: affine gather, rigid gather
    if (s.n_ag > 0) {
      LAUNCH(k_affine_reset, 24 * s.n_ag, s);
      LAUNCH(k_affine_gather, s.n_an, s);
    }
    if (s.n_rg > 0) {
      LAUNCH(k_rigid_reset, 18 * s.n_rg, s);
      LAUNCH(k_rigid_centroid, s.n_rn, s);
      LAUNCH(k_rigid_gather, s.n_rn, s);
    }
    return 0;
  }

  int direction(int it) {
    // gPg, g_p, g_Py, y_p, y_Py were cleared by k_grad_diag_mass
    LAUNCH(k_mask_dk, s.N, s);
    if (s.n_ag > 0) LAUNCH(k_affine_reduce, s.n_ag, s, p.h2);
    if (s.n_rg > 0) LAUNCH(k_rigid_reduce, s.n_rg, s);
    LAUNCH(k_finalize_beta, s.B, s, it, p.cg_guard, (p.ipc_on && pending) ? 1 : 0);
    LAUNCH(k_update_p, s.N, s);
    if (s.n_ag > 0) {
      LAUNCH(k_affine_update, s.n_ag, s, 0);
      LAUNCH(k_affine_expand, s.n_an, s, s.ag_pq, s.p);
    }
    if (s.n_rg > 0) LAUNCH(k_rigid_direction, s.n_rn, s);
    return 0;
  }

  int curvature_caps(int it) {
    LAUNCH(k_pHp_reset, s.B, s, p.alpha_max);
    LAUNCH(k_pHp_mass, s.N, s);
    LAUNCH(k_elastic_pHp, s.T, s.x, s.p, s.tri, s.Bm, s.area, s.mu, s.lam, s.tri2env, s.T, p.h2, 1, s.pHp);
    if (s.n_col > 0) LAUNCH(k_colliders_pHp, s.n_b, s, p);
    if (p.node_on) {
      const int r = node_pass(1, p.h2, s.pHp);
      if (r) return r;
    }
    if (s.n_ag > 0) LAUNCH(k_affine_pHp_energy, s.n_ag, s, 1, p.h2, s.Ener);
    if (p.ipc_on) {
      LAUNCH(k_ipc_begin, s.B, s.B, s.tc, s.alpha, s.clr_ipc, 0);
      int r = ipc_pass(1, s.p, s.grad, s.diag, s.free_, s.Ener, s.clr_ipc, s.alpha, p.h2, p.ipc_mu);
      if (r) return r;
      LAUNCH(k_far_cap, s.B, s.B, s.p_inf, p.ipc_reach, s.alpha);
    }
    if (s.n_sp > 0) LAUNCH(k_springs, s.n_sp, s, 1, p.h2, s.Ener);
    if (s.n_pe > 0) LAUNCH(k_pressure, s.n_pe, s, 1, p.h2, s.Ener);
    if (s.n_td > 0) {
      LAUNCH(k_fill, s.n_td, s.td_dL, s.n_td, 0.0);
      LAUNCH(k_tendon_pHp, s.n_ey, s, p.h2);
      LAUNCH(k_tendon_taut, s.n_td, s, p.h2);
    }
    LAUNCH(k_inversion_alpha, s.T, s.x, s.p, s.tri, s.tri2env, s.T, p.area_shrink, s.alpha);
    if (s.n_col > 0) LAUNCH(k_colliders_ccd, s.n_b, s, p.ccd_eta);
    LAUNCH(k_alpha_finalize, s.B, s, p, it);
    return 0;
  }

  int energy_into(double *E) {
    LAUNCH(k_fill, s.B, E, s.B, 0.0);
    LAUNCH(k_node_energy, s.N, s, p.h2, E);
    LAUNCH(k_elastic_energy, s.T, s.x, s.tri, s.Bm, s.area, s.mu, s.lam, s.tri2env, s.T, p.h2, E);
    if (s.n_col > 0) LAUNCH(k_colliders_energy, s.n_b, s, p, E);
    if (p.node_on) {
      const int r = node_pass(2, p.h2, E);
      if (r) return r;
    }
    if (s.n_ag > 0) LAUNCH(k_affine_pHp_energy, s.n_ag, s, 2, p.h2, E);
    if (p.ipc_on) {
      int r = ipc_pass(2, s.p, s.grad, s.diag, s.free_, E, s.clr_ipc, s.alpha, p.h2, p.ipc_mu);
      if (r) return r;
    }
    if (s.n_sp > 0) LAUNCH(k_springs, s.n_sp, s, 2, p.h2, E);
    if (s.n_pe > 0) LAUNCH(k_pressure, s.n_pe, s, 2, p.h2, E);
    if (s.n_td > 0) {
      LAUNCH(k_fill, s.n_td, s.td_L, s.n_td, 0.0);
      LAUNCH(k_tendon_length, s.n_ey, s);
      LAUNCH(k_tendon_energy, s.n_td, s, p.h2, E);
    }
    return 0;
  }

  int update() {
    if (!p.line_search) {
      LAUNCH(k_update_x, s.N, s, s.x);
      if (s.n_ag > 0) {
        LAUNCH(k_affine_update, s.n_ag, s, 1);
        LAUNCH(k_affine_expand, s.n_an, s, s.ag_q, s.x);
      }
      return 0;
    }
    int r = energy_into(s.Eprev);
    if (r) return r;
    D2F_HIP(hipMemcpyAsync(s.x0, s.x, sizeof(double) * 2 * s.N, hipMemcpyDeviceToDevice, st));
    if (s.n_ag > 0)
      D2F_HIP(hipMemcpyAsync(s.ag_q0, s.ag_q, sizeof(double) * 6 * s.n_ag, hipMemcpyDeviceToDevice, st));
    for (int bt = 0; bt <= p.n_backtrack; ++bt) {
      LAUNCH(k_update_x, s.N, s, s.x0);
      if (s.n_ag > 0) {
        LAUNCH(k_affine_update, s.n_ag, s, 2);
        LAUNCH(k_affine_expand, s.n_an, s, s.ag_q, s.x);
      }
      if (bt == p.n_backtrack) break;
      r = energy_into(s.Ener);
      if (r) return r;
      LAUNCH(k_ls_halve, s.B, s, p.ls_tol);
    }
    return 0;
  }

  int run() {
    LAUNCH(k_init_step, s.N, s, p.dt, p.drag);
    pending = false;
    if (!p.ipc_on) {
      LAUNCH(k_apply_bc, s.n_bc, s);
    } else if (s.n_bc > 0) {
      int r = sweep();
      if (r) return r;
      r = poll_pending(pending);
      if (r) return r;
    }
    LAUNCH(k_reset_step, s.B, s);
    LAUNCH(k_fill, 2 * s.N, s.p, 2 * s.N, 0.0);
    LAUNCH(k_fill, 2 * s.N, s.grad, 2 * s.N, 0.0);
    for (int it = 0; it < p.iter_max; ++it) {
      if (p.ipc_on && pending) {
        int r = sweep();
        if (r) return r;
        LAUNCH(k_bc_wake, s.B, s);
        if (it % 8 == 7) {
          r = poll_pending(pending);
          if (r) return r;
        }
      }
      if (p.conv_check && it && it % p.conv_check == 0) {
        ibuf.resize(s.B);
        D2F_HIP(hipMemcpyAsync(ibuf.data(), s.active, sizeof(int) * s.B, hipMemcpyDeviceToHost, st));
        D2F_HIP(hipStreamSynchronize(st));
        long tot = 0;
        for (int e = 0; e < s.B; ++e) tot += ibuf[e];
        if (tot == 0) break;
      }
      int r = grad_diag();
      if (!r) r = direction(it);
      if (!r) r = curvature_caps(it);
      if (!r) r = update();
      if (r) return r;
    }
    LAUNCH(k_commit_v, s.N, s, p.inv_dt);
    D2F_HIP(hipGetLastError());
    return 0;
  }
};

}  // namespace d2f

using namespace d2f;

// Reaction on the colliders per environment (reference disc_force_kernel, implicit/kernels.py:
// 517-530; sdf_force_kernel, core/sdf.py:185-191): +kappa b' grad phi from every node in the
// barrier; with friction, the reference's friction reactions (disc: +mu f_n t, implicit/
// kernels.py:724-743; other shapes: -mu f_n t, core/sdf.py:259-271), t = v_t / max(|v_t|, epsv).
// index < 0: every collider.
__global__ void k_collider_force(S s, Pm p, int index, int friction, double *F) {
  GRID_LOOP(b, s.n_b) {
    const bool ok = b < s.n_b;
    int e = -1;
    V2<double> f = {0.0, 0.0};
    if (ok) {
      const int n = s.bnd[b];
      e = s.b2e[b];
      const V2<double> x = ld2(s.x, n);
      for (int k = 0; k < s.n_col; ++k) {
        if (index >= 0 && k != index) continue;
        const Collider c = to_collider(s.col[k]);
        if (c.kind == 0) {
          const V2<double> diff = {x.x - c.cx, x.y - c.cy};
          const double dc = len(diff), d = dc - c.r;
          if (d <= 0.0 || d >= p.dhat) continue;
          f = f + (p.kappa * barrier_bp(d, p.dhat) / dc) * diff;
        } else {
          const SdfHit h = sdf_eval(c, x);
          if (h.d <= 0.0 || h.d >= p.dhat) continue;
          f = f + (p.kappa * barrier_bp(h.d, p.dhat)) * h.g;
        }
        if (friction && p.friction) {
          FricFrame fr;
          if (friction_frame(c, x, ld2(s.xn, n), p.dhat, p.kappa, p.inv_dt, fr)) {
            const double sgn = c.kind == 0 ? 1.0 : -1.0;
            f = f + (sgn * c.mu * fr.f_n / fmax(fr.vtm, p.epsv)) * fr.vt;
          }
        }
      }
    }
    env_add(F, 2 * e, f.x, ok);
    env_add(F + 1, 2 * e, f.y, ok);
  }
}

extern "C" {

int d2f_abi_version(void) { return D2F_ABI_VERSION; }
const char *d2f_last_error(void) { return g_err; }
void d2f_struct_sizes(int64_t *out) {
  out[0] = (int64_t)sizeof(d2f_scene);
  out[1] = (int64_t)sizeof(d2f_params);
  out[2] = (int64_t)sizeof(d2f_collider);
  out[3] = (int64_t)sizeof(d2f_minres);
}

int d2f_elastic_grad_diag_f64(const double *x, const int32_t *tri, const double *Bm, const double *area,
                              const double *mu, const double *lam, int32_t T, double h2, int32_t clamp,
                              const d2f_elem_tiles *tiles, double *grad, double *diag, void *stream) {
  D2F_CHECK(T >= 0 && x && tri && Bm && area && mu && lam && grad && diag, "d2f_elastic_grad_diag_f64: bad argument");
  return launch_elastic_grad_diag(x, tri, Bm, area, mu, lam, T, h2, clamp, tiles, grad, diag,
                                  (hipStream_t)stream);
}

int d2f_elastic_pHp_f64(const double *x, const double *p, const int32_t *tri, const double *Bm,
                        const double *area, const double *mu, const double *lam, const int32_t *tri2env,
                        int32_t T, double h2, int32_t clamp, double *pHp, void *stream) {
  D2F_CHECK(T >= 0 && x && p && tri && Bm && area && mu && lam && tri2env && pHp, "d2f_elastic_pHp_f64: bad argument");
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_elastic_pHp, T, x, p, tri, Bm, area, mu, lam, tri2env, T, h2, clamp, pHp);
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_inversion_alpha_f64(const double *x, const double *p, const int32_t *tri, const int32_t *tri2env,
                            int32_t T, double shrink, double *alpha, void *stream) {
  D2F_CHECK(T >= 0 && x && p && tri && tri2env && alpha, "d2f_inversion_alpha_f64: bad argument");
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_inversion_alpha, T, x, p, tri, tri2env, T, shrink, alpha);
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_elastic_energy_f64(const double *x, const int32_t *tri, const double *Bm, const double *area,
                           const double *mu, const double *lam, const int32_t *tri2env, int32_t T,
                           double h2, double *E, void *stream) {
  D2F_CHECK(T >= 0 && x && tri && Bm && area && mu && lam && tri2env && E, "d2f_elastic_energy_f64: bad argument");
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_elastic_energy, T, x, tri, Bm, area, mu, lam, tri2env, T, h2, E);
  D2F_HIP(hipGetLastError());
  return 0;
}

static bool tiles_usable(const d2f_elem_tiles *tiles) {
  return tiles && tiles->node_off && tiles->max_nodes > 0 && tiles->max_nodes <= D2F_ELEM_NODES_MAX;
}

static int launch_hvp_tiled(const double *x, const double *v, const double *Bm, const double *area,
                            const double *mu, const double *lam, int32_t T, double h2, const d2f_elem_tiles *tiles,
                            PressureTiles pt, double *y, hipStream_t st, const char *who) {
  D2F_CHECK(tiles->nodes && tiles->halo_off && tiles->tri_local && tiles->part && tiles->n_hnodes >= 0 &&
                (tiles->n_hnodes == 0 || (tiles->hnode && tiles->hnode_off && tiles->hdest)),
            "%s: incomplete tiles", who);
  const int blocks = (T + D2F_ELEM_CHUNK - 1) / D2F_ELEM_CHUNK;
  const size_t lds = sizeof(double) * 6 * (size_t)tiles->max_nodes;   // at most 48 KiB
  const bool partial = T >= ELEM_PARTIAL_MIN_T;
  if (partial)
    hipLaunchKernelGGL(k_elastic_hvp_tiled<true>, dim3(blocks), dim3(ELEM_TPB), lds, st, x, v, Bm, area, mu, lam, T,
                       h2, *tiles, pt, y);
  else
    hipLaunchKernelGGL(k_elastic_hvp_tiled<false>, dim3(blocks), dim3(ELEM_TPB), lds, st, x, v, Bm, area, mu, lam,
                       T, h2, *tiles, pt, y);
  if (partial && tiles->n_hnodes > 0) {
    const int hb = (int)std::min<long>((tiles->n_hnodes + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_elastic_halo, dim3(hb), dim3(256), 0, st, *tiles, y, (double *)nullptr);
  }
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_elastic_hvp_f64(const double *x, const double *v, const int32_t *tri, const double *Bm,
                        const double *area, const double *mu, const double *lam, int32_t T, double h2,
                        const d2f_elem_tiles *tiles, double *y, void *stream) {
  D2F_CHECK(T >= 0 && x && v && tri && Bm && area && mu && lam && y, "d2f_elastic_hvp_f64: bad argument");
  hipStream_t st = (hipStream_t)stream;
  if (T == 0) return 0;
  if (tiles_usable(tiles))
    return launch_hvp_tiled(x, v, Bm, area, mu, lam, T, h2, tiles, PressureTiles{}, y, st, "d2f_elastic_hvp_f64");
  LAUNCH(k_elastic_hvp, T, x, v, tri, Bm, area, mu, lam, T, h2, y);
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_elastic_pressure_hvp_f64(const double *x, const double *v, const double *Bm, const double *area,
                                 const double *mu, const double *lam, int32_t T, double h2,
                                 const d2f_elem_tiles *tiles, const int32_t *pe_off, const uint32_t *pe_loc,
                                 const int32_t *pe_ch, const double *ch_p, double *y, void *stream) {
  D2F_CHECK(T > 0 && x && v && Bm && area && mu && lam && y && pe_off && pe_loc && pe_ch && ch_p,
            "d2f_elastic_pressure_hvp_f64: bad argument");
  D2F_CHECK(tiles_usable(tiles), "d2f_elastic_pressure_hvp_f64: needs element tiles (max_nodes <= %d)",
            D2F_ELEM_NODES_MAX);
  return launch_hvp_tiled(x, v, Bm, area, mu, lam, T, h2, tiles, PressureTiles{pe_off, pe_loc, pe_ch, ch_p}, y,
                          (hipStream_t)stream, "d2f_elastic_pressure_hvp_f64");
}

int d2f_ipc_min_distance(const d2f_scene *s, const d2f_params *p, void *stream) {
  D2F_CHECK(s && p && p->ipc_on, "d2f_ipc_min_distance: needs a scene with IPC contact");
  hipStream_t st = (hipStream_t)stream;
  // IPCContact.min_distance (reference implicit/ipc_contact.py:441-447): a sweep along a zero
  // direction leaves the smallest pair distance in clr_ipc.
  Plain run(*s, *p, st);
  LAUNCH(k_fill, 2 * s->N, s->x0, 2 * s->N, 0.0);
  LAUNCH(k_fill, s->B, s->q_inf, s->B, 0.0);
  LAUNCH(k_ipc_begin, s->B, s->B, s->tc, s->sweep_beta, s->clr_ipc, 1);
  int r = run.ipc_pass(3, s->x0, s->grad, s->diag, s->free_, s->Ener, s->clr_ipc, s->sweep_beta, p->h2, p->ipc_mu);
  if (r) return r;
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_ipc_pairs(const d2f_scene *s, const d2f_params *p, int32_t *pairs, int32_t cap, void *stream) {
  D2F_CHECK(s && p && p->ipc_on && pairs && cap >= 0, "d2f_ipc_pairs: needs a scene with IPC contact");
  hipStream_t st = (hipStream_t)stream;
  D2F_HIP(hipMemsetAsync(pairs, 0, sizeof(int32_t), st));
  Plain run(*s, *p, st);
  run.pc.plist = pairs;
  run.pc.pcap = cap;
  int r = run.ipc_pass(4, s->x0, s->grad, s->diag, s->free_, s->Ener, s->clr_ipc, s->sweep_beta, p->h2, p->ipc_mu);
  if (r) return r;
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_step_plain(const d2f_scene *s, const d2f_params *p, void *stream) {
  D2F_CHECK(s && p, "d2f_step_plain: null scene or params");
  D2F_CHECK(!p->node_on || (!p->ipc_on && s->excl_bits && s->excl_bits_off && p->node_r > 0.0),
            "d2f_step_plain: node contact needs excl_bits, excl_bits_off, node_r > 0 and no IPC contact");
  D2F_CHECK(s->B > 0 && s->N > 0 && s->T >= 0 && s->n_col >= 0 && s->n_col <= D2F_MAX_COLLIDERS,
            "d2f_step_plain: bad sizes (B=%d N=%d T=%d n_col=%d)", s->B, s->N, s->T, s->n_col);
  D2F_CHECK(p->iter_max >= 0 && p->dt > 0.0, "d2f_step_plain: bad params");
  Plain run(*s, *p, (hipStream_t)stream);
  return run.run();
}

int d2f_apply_activation_f64(const int32_t *elem, const int32_t *ereg, int32_t n_el, const double *Bm0,
                             const double *area0, const double *Ainv, const double *detA,
                             const d2f_elem_consts *out, void *stream) {
  D2F_CHECK(n_el >= 0 && out, "d2f_apply_activation_f64: bad argument");
  if (n_el == 0) return 0;
  D2F_CHECK(elem && ereg && Bm0 && area0 && Ainv && detA && out->Bm && out->area,
            "d2f_apply_activation_f64: null pointer");
  D2F_CHECK(!out->sc_pos || (out->mu && out->lam && out->sc_Bm && out->sc_Amu && out->sc_Alam && out->sc_Bm32 &&
                             out->sc_Amu32 && out->sc_Alam32),
            "d2f_apply_activation_f64: scene layout incomplete");
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_apply_activation, n_el, elem, ereg, n_el, Bm0, area0, Ainv, detA, *out);
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_collider_force(const d2f_scene *s, const d2f_params *p, int32_t index, int32_t friction, double *F,
                       void *stream) {
  D2F_CHECK(s && p && F, "d2f_collider_force: null argument");
  D2F_CHECK(index < s->n_col, "d2f_collider_force: collider %d of %d", index, s->n_col);
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_fill, 2 * s->B, F, 2 * s->B, 0.0);
  if (s->n_b > 0 && s->n_col > 0) LAUNCH(k_collider_force, s->n_b, *s, *p, index, friction, F);
  D2F_HIP(hipGetLastError());
  return 0;
}

int d2f_contact_force(const d2f_scene *s, const d2f_params *p, int32_t colliders, int32_t bodies,
                      int32_t friction, const uint8_t *all_free, double *g, double *dg, void *stream) {
  D2F_CHECK(s && p && all_free && g && dg, "d2f_contact_force: null argument");
  hipStream_t st = (hipStream_t)stream;
  LAUNCH(k_fill, 2 * s->N, g, 2 * s->N, 0.0);
  LAUNCH(k_fill, 2 * s->N, dg, 2 * s->N, 0.0);
  // the gradient passes with h^2 = 1, nothing masked, into g (reference implicit/solver.py:554-602)
  S s1 = *s;
  s1.grad = g;
  s1.diag = dg;
  s1.free_ = all_free;
  Pm p1 = *p;
  p1.h2 = 1.0;
  p1.friction = friction ? p->friction : 0;
  if (colliders && s->n_b > 0 && s->n_col > 0) LAUNCH(k_colliders_grad_diag, s->n_b, s1, p1);
  if (bodies && p->ipc_on && s->n_e > 0) {
    Pm p2 = p1;
    if (!friction) {
      p2.ipc_mu = 0.0;
      p2.ipc_per_body = 0;
    }
    Plain run(s1, p2, st);
    const int r = run.ipc_pass(0, s->p, g, dg, all_free, s->Ener, s->clr_min, s->alpha, 1.0, p2.ipc_mu);
    if (r) return r;
  }
  if (bodies && p->node_on) {  // reference implicit/solver.py:591-601: penalty, then pair friction
    Pm p2 = p1;
    p2.node_fric = friction ? p->node_fric : 0;
    Plain run(s1, p2, st);
    const int r = run.node_pass(0, 1.0, nullptr);
    if (r) return r;
  }
  D2F_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
