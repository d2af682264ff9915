// adjoint.hip — the tangent pieces of the implicit adjoint (reference implicit/adjoint.py):
// collider barrier Hessian-vector products, the collider gradient/diagonal on its own (the
// Jacobi preconditioner of the adjoint solve) and dJ/drho per element. All float64, one thread
// per boundary node or per element; the element Hessian-vector product is d2f_elastic_hvp_f64.
#include "common.h"
#include "d2f_math.h"
#include "node_contact.h"

namespace d2f {

struct ColSet {
  int n;
  Collider c[D2F_MAX_COLLIDERS];
};

// which: 0 = the discs (kind 0), 1 = the other shapes through their level sets, 2 = all
__device__ __forceinline__ bool col_selected(int kind, int which) {
  return which == 2 || (which == 0 ? kind == 0 : kind != 0);
}

// y += h2 kappa (b''(d) (g.v) g + b'(d) (Hphi v)) for 0 < d < dhat, the contact gate frozen at x
// (reference chvp_disc_kernel, implicit/adjoint.py:89-99; chvp_sdf_kernel, :103-112). UNCLAMPED.
__global__ void k_collider_hvp(const double *x, const double *v, const int32_t *bnd, int n_b, ColSet cs,
                               int which, double dhat, double kappa, double h2, double *y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b) return;
  const int n = bnd[b];
  const V2<double> xn = ld2(x, n), vn = ld2(v, n);
  V2<double> acc = {0.0, 0.0};
  bool any = false;
  for (int k = 0; k < cs.n; ++k) {
    const Collider &c = cs.c[k];
    if (!col_selected(c.kind, which)) continue;
    if (c.kind == 0) {
      const V2<double> diff = {xn.x - c.cx, xn.y - c.cy};
      const double dc = len(diff), d = dc - c.r;
      if (d <= 0.0 || d >= dhat) continue;
      const V2<double> nh = {diff.x / dc, diff.y / dc};
      const double pn = dot(vn, nh);
      const V2<double> pt = vn - pn * nh;
      acc = acc + kappa * (barrier_bpp(d, dhat) * pn * nh + (barrier_bp(d, dhat) / dc) * pt);
    } else {
      const SdfHit s = sdf_eval(c, xn);
      if (s.d <= 0.0 || s.d >= dhat) continue;
      const double gv = dot(s.g, vn);
      const V2<double> Hv = {s.H.a * vn.x + s.H.b * vn.y, s.H.c * vn.x + s.H.d * vn.y};
      acc = acc + kappa * (barrier_bpp(s.d, dhat) * gv * s.g + barrier_bp(s.d, dhat) * Hv);
    }
    any = true;
  }
  if (any) st2(y, n, ld2(y, n) + h2 * acc);
}

// grad += h2 kappa b' grad phi, diag += h2 (clamped) curvature of the selected colliders
// (reference grad_diag_disc_kernel, implicit/kernels.py:132-157; grad_diag_sdf_kernel,
// core/sdf.py:141-154) — the contact part of the adjoint's Jacobi preconditioner.
__global__ void k_collider_grad_diag(const double *x, const int32_t *bnd, int n_b, ColSet cs, int which,
                                     double dhat, double kappa, double h2, int clamp, double *g,
                                     double *dg) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b) return;
  const int n = bnd[b];
  const V2<double> xn = ld2(x, n);
  V2<double> ga = ld2(g, n), da = ld2(dg, n);
  for (int k = 0; k < cs.n; ++k) {
    if (!col_selected(cs.c[k].kind, which)) continue;
    V2<double> gg, dd;
    if (collider_grad_diag(cs.c[k], xn, dhat, kappa, h2, clamp, gg, dd)) {
      ga = ga + gg;
      da = da + dd;
    }
  }
  st2(g, n, ga);
  st2(dg, n, da);
}

// out[t] += -(1 - Emin) h2 lam . f_unit_t(x): the element's full-density elastic force paired
// with the adjoint vector (reference drho_kernel, implicit/adjoint.py:165-178).
__global__ void k_elastic_drho(const double *x, const double *lam, const int32_t *tri, const double *Bm,
                               const double *area, const double *mu0, const double *lam0, int T,
                               double h2, double Emin, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
  const V2<double> p0 = ld2(x, i0), p1 = ld2(x, i1), p2 = ld2(x, i2);
  const double4 bb = reinterpret_cast<const double4 *>(Bm)[t];
  const M2<double> B = {bb.x, bb.y, bb.z, bb.w};
  const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
  const NH<double> h = nh_eval(mat_mul(Ds, B), mu0[t], lam0[t]);
  V2<double> g0, g1, g2;
  nh_grad(h, B, mu0[t], area[t], g0, g1, g2);
  const double s = dot(ld2(lam, i0), g0) + dot(ld2(lam, i1), g1) + dot(ld2(lam, i2), g2);
  out[t] += -(1.0 - Emin) * h2 * s;
}

// ---- pressure (not in the reference's adjoint, which refuses pressure chambers): the gradient
// of E = -p A(x) puts -1/2 h2 p J(x_b - x_a) on both ends of every chamber edge, J(u) =
// (u_y, -u_x) (reference implicit/actuation.py:126-151); it is linear in x.
// y += h2 H_pressure v: -1/2 h2 p J(v_b - v_a) on both ends.
__global__ void k_pressure_hvp(const double *v, const int32_t *pe_a, const int32_t *pe_b, const int32_t *pe_ch,
                               const double *ch_p, int n_pe, double h2, double *y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_pe) return;
  const int a = pe_a[k], b = pe_b[k];
  const V2<double> u = ld2(v, b) - ld2(v, a);
  const double c = -0.5 * h2 * ch_p[pe_ch[k]];
  const V2<double> w = {c * u.y, -c * u.x};
  atomic_add2(y, a, w);
  atomic_add2(y, b, w);
}

// ---- IPC contact between bodies (not in the reference's adjoint, which refuses IPCContact).
// The barrier terms of one (edge, node) pair are those the gradient pass evaluates (reference
// implicit/ipc_contact.py:131-296): the interior term when the foot is inside the edge and
// d < dhat, and an end term per edge end whose vertex-correction coefficient is non-zero and
// whose distance to the node is below dhat. The set and the coefficients are frozen at x.
struct PairGeom {
  int p, a, b;
  V2<double> xp, xa, xb;
  double tt;        // foot parameter along a -> b
  double d;         // distance to the segment
  double coef_a, coef_b;
};

__device__ __forceinline__ PairGeom pair_geom(const double *x, const int32_t *pairs, int i, const int32_t *ea,
                                              const int32_t *eb, const double *ca, const double *cb) {
  PairGeom g;
  const int k = pairs[2 + 2 * i];
  g.p = pairs[3 + 2 * i];
  g.a = ea[k];
  g.b = eb[k];
  g.xp = ld2(x, g.p);
  g.xa = ld2(x, g.a);
  g.xb = ld2(x, g.b);
  const V2<double> ab = g.xb - g.xa, w = g.xp - g.xa;
  g.tt = dot(w, ab) / fmax(dot(ab, ab), 1.0e-30);
  V2<double> dv = w - g.tt * ab;
  if (g.tt <= 0.0) dv = w;
  else if (g.tt >= 1.0) dv = g.xp - g.xb;
  g.d = len(dv);
  g.coef_a = -ca[k] + (g.tt <= 0.0 ? 1.0 : 0.0);
  g.coef_b = -cb[k] + (g.tt >= 1.0 ? 1.0 : 0.0);
  return g;
}

__device__ __forceinline__ V2<double> perp(V2<double> u) { return {-u.y, u.x}; }

// y += h2 (Hessian of the pair's barrier energy) v, UNCLAMPED, friction left out. Interior term:
// d = |s|, s = (e x w) / L with e = x_b - x_a, w = x_p - x_a, L = |e|; grad s is perp(e) / L on
// p and -perp(w) / L - s e / L^2 on b (minus their sum on a), and the product takes the
// directional derivative of grad s along v: H v = kappa (b'' (grad s . v) grad s + b' sign(s)
// D(grad s)[v]). End terms: d = |x_p - x_end|, Hess d = (I - n n^T) / d.
__global__ void k_ipc_hvp(const double *x, const double *v, const int32_t *pairs, int n_pairs, const int32_t *ea,
                          const int32_t *eb, const double *ca, const double *cb, double dhat, double kappa,
                          double h2, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs) return;
  const PairGeom g = pair_geom(x, pairs, i, ea, eb, ca, cb);
  const double dlo = 1.0e-9 * dhat;
  const V2<double> vp = ld2(v, g.p), va = ld2(v, g.a), vb = ld2(v, g.b);
  V2<double> yp = {0.0, 0.0}, ya = {0.0, 0.0}, yb = {0.0, 0.0};
  if (g.tt > 0.0 && g.tt < 1.0 && g.d < dhat) {
    const V2<double> e = g.xb - g.xa, w = g.xp - g.xa, de = vb - va, dw = vp - va;
    const double L2 = fmax(dot(e, e), 1.0e-30), L = sqrt(L2), L3 = L * L2;
    const double c = cross(e, w);
    const double d = fmax(fabs(c) / L, dlo);
    const double sg = c < 0.0 ? -1.0 : 1.0;
    const V2<double> gw = (1.0 / L) * perp(e);
    const V2<double> ge = (-1.0 / L) * perp(w) - (c / L3) * e;
    const double sv = dot(gw, dw) + dot(ge, de);
    const double eL = dot(e, de);
    const double dc = cross(de, w) + cross(e, dw);
    const V2<double> dgw = (1.0 / L) * perp(de) - (eL / L3) * perp(e);
    const V2<double> dge = (-1.0 / L) * perp(dw) + (eL / L3) * perp(w) - (dc / L3) * e - (c / L3) * de +
                           (3.0 * c * eL / (L2 * L3)) * e;
    const double k = kappa * h2;
    const double bp = barrier_bp(d, dhat) * sg, bpp = barrier_bpp(d, dhat) * sv;
    const V2<double> Hw = k * (bpp * gw + bp * dgw);
    const V2<double> He = k * (bpp * ge + bp * dge);
    yp = yp + Hw;
    yb = yb + He;
    ya = ya - (Hw + He);
  }
  for (int end = 0; end < 2; ++end) {
    const double coef = end == 0 ? g.coef_a : g.coef_b;
    const V2<double> w = end == 0 ? g.xp - g.xa : g.xp - g.xb;
    const double dd = len(w);
    if (coef == 0.0 || !(dd < dhat)) continue;
    const double d = fmax(dd, dlo);
    const V2<double> n = (1.0 / d) * w;
    const V2<double> dq = vp - (end == 0 ? va : vb);
    const double dn = dot(dq, n);
    const V2<double> H = (coef * kappa * h2) *
                         ((barrier_bpp(d, dhat) * dn) * n + (barrier_bp(d, dhat) / d) * (dq - dn * n));
    yp = yp + H;
    if (end == 0) ya = ya - H;
    else yb = yb - H;
  }
  atomic_add2(y, g.p, yp);
  atomic_add2(y, g.a, ya);
  atomic_add2(y, g.b, yb);
}

// grad += h2 kappa b' grad d and diag += h2 kappa b'' (w grad d)^2 per term (the gradient pass
// with friction off, nothing masked): the contact part of the adjoint's Jacobi preconditioner.
__global__ void k_ipc_grad_diag(const double *x, const int32_t *pairs, int n_pairs, const int32_t *ea,
                                const int32_t *eb, const double *ca, const double *cb, double dhat, double kappa,
                                double h2, double *gout, double *dout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs) return;
  const PairGeom g = pair_geom(x, pairs, i, ea, eb, ca, cb);
  const double dlo = 1.0e-9 * dhat;
  auto term = [&](double coef, double wa, double wb, double d, V2<double> n) {
    const V2<double> f = (coef * h2 * kappa * barrier_bp(d, dhat)) * n;
    const double c2 = coef * h2 * kappa * barrier_bpp(d, dhat);
    const V2<double> hd = {c2 * n.x * n.x, c2 * n.y * n.y};
    atomic_add2(gout, g.p, f);
    atomic_add2(dout, g.p, hd);
    if (wa != 0.0) {
      atomic_add2(gout, g.a, (-wa) * f);
      atomic_add2(dout, g.a, (wa * wa) * hd);
    }
    if (wb != 0.0) {
      atomic_add2(gout, g.b, (-wb) * f);
      atomic_add2(dout, g.b, (wb * wb) * hd);
    }
  };
  if (g.tt > 0.0 && g.tt < 1.0 && g.d < dhat) {
    const V2<double> ab = g.xb - g.xa, w = g.xp - g.xa;
    const double d = fmax(g.d, dlo);
    term(1.0, 1.0 - g.tt, g.tt, d, (1.0 / d) * (w - g.tt * ab));
  }
  const V2<double> wa = g.xp - g.xa, wb = g.xp - g.xb;
  const double da = len(wa), db = len(wb);
  if (g.coef_a != 0.0 && da < dhat) term(g.coef_a, 1.0, 0.0, fmax(da, dlo), (1.0 / fmax(da, dlo)) * wa);
  if (g.coef_b != 0.0 && db < dhat) term(g.coef_b, 0.0, 1.0, fmax(db, dlo), (1.0 / fmax(db, dlo)) * wb);
}

// ---- affine bodies in the adjoint (not in the reference's adjoint, which refuses
// AffineBodies). A slaved node follows x_i = J_i q, J_i = diag(w_i^T, w_i^T), w_i = (1, X_i - c),
// q = (t_x, A11, A12, t_y, A21, A22) (reference implicit/affine.py:216-224). The adjoint solves
// in (free soft nodes, q per group); a reduced vector keeps the soft nodes' entries in place and
// group g's six q entries in the (x, y) slots of three of its nodes, slot[3g .. 3g + 2].
// vf = P vr: soft nodes copied, slaved nodes J_i q.
__global__ void k_affine_expand(const double *vr, const int32_t *slave_of, const int32_t *an_grp,
                                const double *an_w, const int32_t *slot, int N, double *vf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int j = slave_of[i];
  if (j < 0) {
    st2(vf, i, ld2(vr, i));
    return;
  }
  const int g = an_grp[j];
  const V2<double> q01 = ld2(vr, slot[3 * g]), q23 = ld2(vr, slot[3 * g + 1]), q45 = ld2(vr, slot[3 * g + 2]);
  const double w0 = an_w[3 * j], w1 = an_w[3 * j + 1], w2 = an_w[3 * j + 2];
  st2(vf, i, V2<double>{w0 * q01.x + w1 * q01.y + w2 * q23.x, w0 * q23.y + w1 * q45.x + w2 * q45.y});
}

// yr = y on soft nodes, 0 on slaved ones (the groups' slots are written by k_affine_reduce_groups).
__global__ void k_affine_reduce_nodes(const double *yf, const int32_t *slave_of, int N, double *yr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  st2(yr, i, slave_of[i] < 0 ? ld2(yf, i) : V2<double>{0.0, 0.0});
}

constexpr int AR_TPB = 256;

// Group g's slots = sum over its nodes of (w_i u_x, w_i u_y), u = yf + mass vf when vf is given
// (the slaved nodes' inertia, which the Krylov solve's diagonal mass leaves out), or of
// (w_i^2 y_x, w_i^2 y_y) with square (the diagonal of J^T diag(y) J). One workgroup per group.
__global__ __launch_bounds__(AR_TPB) void k_affine_reduce_groups(const double *yf, const double *vf,
                                                                 const double *mass, const int32_t *an_node,
                                                                 const double *an_w, const int32_t *an_first,
                                                                 const int32_t *slot, int square, double *yr) {
  __shared__ double part[AR_TPB / WAVE][6];
  const int g = blockIdx.x;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = an_first[g] + threadIdx.x; j < an_first[g + 1]; j += AR_TPB) {
    const int i = an_node[j];
    V2<double> u = ld2(yf, i);
    if (vf) u = u + mass[i] * ld2(vf, i);
    double w[3] = {an_w[3 * j], an_w[3 * j + 1], an_w[3 * j + 2]};
    if (square)
      for (int k = 0; k < 3; ++k) w[k] *= w[k];
    for (int k = 0; k < 3; ++k) {
      s[k] += w[k] * u.x;
      s[3 + k] += w[k] * u.y;
    }
  }
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  for (int k = 0; k < 6; ++k) {
    double t = s[k];
    for (int o = WAVE / 2; o > 0; o >>= 1) t += __shfl_down(t, o, WAVE);
    if (lane == 0) part[wv][k] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double q[6];
    for (int k = 0; k < 6; ++k) {
      q[k] = 0.0;
      for (int w = 0; w < AR_TPB / WAVE; ++w) q[k] += part[w][k];
    }
    st2(yr, slot[3 * g], V2<double>{q[0], q[1]});
    st2(yr, slot[3 * g + 1], V2<double>{q[2], q[3]});
    st2(yr, slot[3 * g + 2], V2<double>{q[4], q[5]});
  }
}

// out[c] += -lam . dR/dp_c = 1/2 h2 sum over chamber c's edges of (lam_a + lam_b) . J(x_b - x_a).
__global__ void k_pressure_dp(const double *x, const double *lam, const int32_t *pe_a, const int32_t *pe_b,
                              const int32_t *pe_ch, int n_pe, double h2, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_pe) return;
  const int a = pe_a[k], b = pe_b[k];
  const V2<double> t = ld2(x, b) - ld2(x, a);
  const V2<double> l = ld2(lam, a) + ld2(lam, b);
  atomicAdd(out + pe_ch[k], 0.5 * h2 * (l.x * t.y - l.y * t.x));
}

// ---- MINRES with per-scene Krylov scalars (Paige and Saunders, SIAM J. Numer. Anal. 12,
// 1975), for the adjoint solve H lam = b of every scene at once: scene e's degrees of freedom
// are [2 node_start[e], 2 node_start[e+1]); each scene has its own Lanczos process, so scenes
// do not couple through shared scalars. One workgroup per scene; its scalars and status stay
// on the device between launches, so an iteration needs no host round trip. The operator is
// applied outside: the caller fills Kv = h2 K v (the Hessian-vector kernels, v as left by the
// previous launch), and d2f_minres_step_f64 forms H v = fmask (mass v + Kv).
constexpr int MR_TPB = 256;
constexpr int MR_NS = 8;  // beta1, oldb, beta, dbar, epsln, phibar, cs, sn

__device__ __forceinline__ double mr_block_sum(double v, double *red) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < MR_TPB / WAVE; ++k) t += red[k];
  return t;
}

__global__ __launch_bounds__(MR_TPB) void k_minres_begin(d2f_minres m) {
  __shared__ double red[MR_TPB / WAVE];
  const int e = blockIdx.x;
  const int d0 = 2 * m.node_start[e], d1 = 2 * m.node_start[e + 1];
  double s = 0.0;
  for (int i = d0 + threadIdx.x; i < d1; i += MR_TPB) {
    const double r = m.b[i] * m.fmask[i];
    const double y = m.dinv[i] * r;
    m.r1[i] = r;
    m.r2[i] = r;
    m.Kv[i] = y;
    m.x[i] = 0.0;
    m.w1[i] = 0.0;
    m.w2[i] = 0.0;
    s += r * y;
  }
  const double bb = mr_block_sum(s, red);
  const double beta1 = bb > 0.0 ? sqrt(bb) : 0.0;
  for (int i = d0 + threadIdx.x; i < d1; i += MR_TPB) {
    m.v[i] = beta1 > 0.0 ? m.Kv[i] / beta1 : 0.0;
    m.Kv[i] = 0.0;                    // cleared for the first product
  }
  if (threadIdx.x == 0) {
    double *q = m.scal + MR_NS * e;
    q[0] = beta1;
    q[1] = 0.0;
    q[2] = beta1;
    q[3] = 0.0;
    q[4] = 0.0;
    q[5] = beta1;
    q[6] = -1.0;
    q[7] = 0.0;
    m.status[2 * e] = bb < 0.0 ? -1 : (beta1 == 0.0 ? 1 : 0);
    m.status[2 * e + 1] = 0;
  }
}

// Scenes of up to MR_TPB * MR_K degrees of freedom keep each thread's entries of v and of the
// Lanczos vector in registers between the step's three passes, and each pass loads all of a
// thread's entries before it computes or stores any (indices past the scene are clamped to its
// first entry and their values unused); larger scenes pass them through Kv in memory. Both give
// the same values: every entry's arithmetic and each thread's summation order are the same.
constexpr int MR_K = 16;

__global__ __launch_bounds__(MR_TPB) void k_minres_step(d2f_minres m) {
  __shared__ double red[MR_TPB / WAVE];
  const int e = blockIdx.x;
  if (m.status[2 * e] != 0) return;
  const int d0 = 2 * m.node_start[e], d1 = 2 * m.node_start[e + 1];
  double *q = m.scal + MR_NS * e;
  const double beta1 = q[0], oldb = q[1], beta = q[2], dbar = q[3], epsln = q[4], phibar = q[5], cs = q[6],
               sn = q[7];
  const int it = m.status[2 * e + 1] + 1;
  const double c1 = it >= 2 ? beta / oldb : 0.0;
  const bool regs = d1 - d0 <= MR_TPB * MR_K;
  double vk[MR_K], yk[MR_K];
  // y = H v - (beta / oldb) r1; alfa = v . y
  double s = 0.0;
  if (regs) {
    double kv[MR_K], r1[MR_K], f[MR_K], ms[MR_K];
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      const int i = d0 + threadIdx.x + k * MR_TPB, j = i < d1 ? i : d0;
      vk[k] = m.v[j];
      kv[k] = m.Kv[j];
      r1[k] = m.r1[j];
      f[k] = m.fmask[j];
      ms[k] = m.mass[j >> 1];
    }
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      yk[k] = f[k] * (ms[k] * vk[k] + kv[k]) - c1 * r1[k];
      if (d0 + (int)threadIdx.x + k * MR_TPB < d1) s += vk[k] * yk[k];
    }
  } else {
    for (int i = d0 + threadIdx.x; i < d1; i += MR_TPB) {
      const double y = m.fmask[i] * (m.mass[i >> 1] * m.v[i] + m.Kv[i]) - c1 * m.r1[i];
      m.Kv[i] = y;
      s += m.v[i] * y;
    }
  }
  const double alfa = mr_block_sum(s, red);
  // y -= (alfa / beta) r2; r1 <- r2; r2 <- y; y <- dinv r2; beta' = sqrt(r2 . y)
  s = 0.0;
  const double c2 = alfa / beta;
  if (regs) {
    double r2[MR_K], dv[MR_K];
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      const int i = d0 + threadIdx.x + k * MR_TPB, j = i < d1 ? i : d0;
      r2[k] = m.r2[j];
      dv[k] = m.dinv[j];
    }
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      const int i = d0 + threadIdx.x + k * MR_TPB;
      const double y = yk[k] - c2 * r2[k];
      yk[k] = dv[k] * y;
      if (i < d1) {
        m.r1[i] = r2[k];
        m.r2[i] = y;
        s += y * yk[k];
      }
    }
  } else {
    for (int i = d0 + threadIdx.x; i < d1; i += MR_TPB) {
      const double y = m.Kv[i] - c2 * m.r2[i];
      m.r1[i] = m.r2[i];
      m.r2[i] = y;
      const double z = m.dinv[i] * y;
      m.Kv[i] = z;
      s += y * z;
    }
  }
  const double bb = mr_block_sum(s, red);
  if (bb < 0.0) {
    if (threadIdx.x == 0) {
      m.status[2 * e] = -1;
      m.status[2 * e + 1] = it;
    }
    return;
  }
  const double nb = sqrt(bb);
  const double oldeps = epsln;
  const double delta = cs * dbar + sn * alfa;
  const double gbar = sn * dbar - cs * alfa;
  const double eps2 = sn * nb;
  const double dbar2 = -cs * nb;
  const double gamma = fmax(sqrt(gbar * gbar + nb * nb), 1.0e-300);
  const double cs2 = gbar / gamma, sn2 = nb / gamma;
  const double phi = cs2 * phibar;
  const double phibar2 = sn2 * phibar;
  // w <- (v - oldeps w1 - delta w2) / gamma, (w1, w2) <- (w2, w); x += phi w; v <- y / beta';
  // Kv cleared for the next product
  if (regs) {
    double w1[MR_K], w2[MR_K], xx[MR_K];
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      const int i = d0 + threadIdx.x + k * MR_TPB, j = i < d1 ? i : d0;
      w1[k] = m.w1[j];
      w2[k] = m.w2[j];
      xx[k] = m.x[j];
    }
#pragma unroll
    for (int k = 0; k < MR_K; ++k) {
      const int i = d0 + threadIdx.x + k * MR_TPB;
      if (i < d1) {
        const double w = (vk[k] - oldeps * w1[k] - delta * w2[k]) / gamma;
        m.w1[i] = w2[k];
        m.w2[i] = w;
        m.x[i] = xx[k] + phi * w;
        m.v[i] = nb > 0.0 ? yk[k] / nb : 0.0;
        m.Kv[i] = 0.0;
      }
    }
  } else {
    for (int i = d0 + threadIdx.x; i < d1; i += MR_TPB) {
      const double w = (m.v[i] - oldeps * m.w1[i] - delta * m.w2[i]) / gamma;
      m.w1[i] = m.w2[i];
      m.w2[i] = w;
      m.x[i] += phi * w;
      m.v[i] = nb > 0.0 ? m.Kv[i] / nb : 0.0;
      m.Kv[i] = 0.0;
    }
  }
  if (threadIdx.x == 0) {
    q[1] = beta;
    q[2] = nb;
    q[3] = dbar2;
    q[4] = eps2;
    q[5] = phibar2;
    q[6] = cs2;
    q[7] = sn2;
    m.status[2 * e] = (phibar2 <= m.rtol * beta1 || nb == 0.0) ? 1 : (it >= m.maxiter ? 2 : 0);
    m.status[2 * e + 1] = it;
  }
}

static ColSet colset(const d2f_collider *cols, int n_col) {
  ColSet cs{};
  cs.n = n_col;
  for (int k = 0; k < n_col; ++k) cs.c[k] = to_collider(cols[k]);
  return cs;
}

}  // namespace d2f

using namespace d2f;

extern "C" int d2f_collider_hvp_f64(const double *x, const double *v, const int32_t *bnd, int32_t n_b,
                                    const d2f_collider *cols, int32_t n_col, int32_t which, double dhat,
                                    double kappa, double h2, double *y, void *stream) {
  D2F_CHECK(n_b >= 0 && n_col >= 0 && n_col <= D2F_MAX_COLLIDERS && which >= 0 && which <= 2,
            "d2f_collider_hvp_f64: bad argument");
  D2F_CHECK(n_b == 0 || n_col == 0 || (x && v && bnd && cols && y), "d2f_collider_hvp_f64: null pointer");
  if (n_b == 0 || n_col == 0) return 0;
  hipLaunchKernelGGL(k_collider_hvp, dim3((n_b + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, v, bnd,
                     n_b, colset(cols, n_col), which, dhat, kappa, h2, y);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_collider_grad_diag_f64(const double *x, const int32_t *bnd, int32_t n_b,
                                          const d2f_collider *cols, int32_t n_col, int32_t which,
                                          double dhat, double kappa, double h2, int32_t clamp, double *g,
                                          double *dg, void *stream) {
  D2F_CHECK(n_b >= 0 && n_col >= 0 && n_col <= D2F_MAX_COLLIDERS && which >= 0 && which <= 2,
            "d2f_collider_grad_diag_f64: bad argument");
  D2F_CHECK(n_b == 0 || n_col == 0 || (x && bnd && cols && g && dg), "d2f_collider_grad_diag_f64: null pointer");
  if (n_b == 0 || n_col == 0) return 0;
  hipLaunchKernelGGL(k_collider_grad_diag, dim3((n_b + 255) / 256), dim3(256), 0, (hipStream_t)stream, x,
                     bnd, n_b, colset(cols, n_col), which, dhat, kappa, h2, clamp, g, dg);
  D2F_HIP(hipGetLastError());
  return 0;
}

// The "node" contact's penalty on its own (reference implicit/kernels.py:772-808 as
// implicit/adjoint.py:319-336 calls it, and implicit/adjoint.py:115-143).
static NodeArgs node_args(const double *x, const double *v, const int32_t *bnd, const int32_t *bnd_start,
                          int32_t B, const uint32_t *excl_bits, const int64_t *excl_bits_off, double radius,
                          double stiffness, double h2, int32_t clamp, double *g, double *dg) {
  NodeArgs a{};
  a.x = x;
  a.pdir = v;
  a.bnd = bnd;
  a.bnd_start = bnd_start;
  a.B = B;
  a.max_nb = 0;  // sizes unknown on the host: partners read from global memory
  a.excl_bits = excl_bits;
  a.excl_bits_off = excl_bits_off;
  a.k = stiffness;
  a.r = radius;
  a.h2 = h2;
  a.clamp = clamp;
  a.g = g;
  a.dg = dg;
  return a;
}

extern "C" int d2f_node_grad_diag_f64(const double *x, const int32_t *bnd, const int32_t *bnd_start, int32_t B,
                                      const uint32_t *excl_bits, const int64_t *excl_bits_off, double radius,
                                      double stiffness, double h2, int32_t clamp, double *g, double *dg,
                                      void *stream) {
  D2F_CHECK(B >= 0 && radius >= 0.0, "d2f_node_grad_diag_f64: bad argument");
  D2F_CHECK(B == 0 || (x && bnd && bnd_start && excl_bits && excl_bits_off && g && dg),
            "d2f_node_grad_diag_f64: null pointer");
  return launch_node_env(0, node_args(x, nullptr, bnd, bnd_start, B, excl_bits, excl_bits_off, radius, stiffness, h2,
                                      clamp, g, dg),
                         (hipStream_t)stream);
}

extern "C" int d2f_node_hvp_f64(const double *x, const double *v, const int32_t *bnd, const int32_t *bnd_start,
                                int32_t B, const uint32_t *excl_bits, const int64_t *excl_bits_off, double radius,
                                double stiffness, double h2, double *y, void *stream) {
  D2F_CHECK(B >= 0 && radius >= 0.0, "d2f_node_hvp_f64: bad argument");
  D2F_CHECK(B == 0 || (x && v && bnd && bnd_start && excl_bits && excl_bits_off && y), "d2f_node_hvp_f64: null pointer");
  return launch_node_env(3, node_args(x, v, bnd, bnd_start, B, excl_bits, excl_bits_off, radius, stiffness, h2, 0, y,
                                      nullptr),
                         (hipStream_t)stream);
}

extern "C" int d2f_elastic_drho_f64(const double *x, const double *lam, const int32_t *tri, const double *Bm,
                                    const double *area, const double *mu0, const double *lam0, int32_t T,
                                    double h2, double Emin, double *out, void *stream) {
  D2F_CHECK(T >= 0, "d2f_elastic_drho_f64: bad argument");
  D2F_CHECK(T == 0 || (x && lam && tri && Bm && area && mu0 && lam0 && out), "d2f_elastic_drho_f64: null pointer");
  if (T == 0) return 0;
  hipLaunchKernelGGL(k_elastic_drho, dim3((T + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, lam, tri,
                     Bm, area, mu0, lam0, T, h2, Emin, out);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_pressure_hvp_f64(const double *v, const int32_t *pe_a, const int32_t *pe_b, const int32_t *pe_ch,
                                    const double *ch_p, int32_t n_pe, double h2, double *y, void *stream) {
  D2F_CHECK(n_pe >= 0, "d2f_pressure_hvp_f64: bad argument");
  D2F_CHECK(n_pe == 0 || (v && pe_a && pe_b && pe_ch && ch_p && y), "d2f_pressure_hvp_f64: null pointer");
  if (n_pe == 0) return 0;
  hipLaunchKernelGGL(k_pressure_hvp, dim3((n_pe + 255) / 256), dim3(256), 0, (hipStream_t)stream, v, pe_a, pe_b,
                     pe_ch, ch_p, n_pe, h2, y);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_pressure_dp_f64(const double *x, const double *lam, const int32_t *pe_a, const int32_t *pe_b,
                                   const int32_t *pe_ch, int32_t n_pe, double h2, double *out, void *stream) {
  D2F_CHECK(n_pe >= 0, "d2f_pressure_dp_f64: bad argument");
  D2F_CHECK(n_pe == 0 || (x && lam && pe_a && pe_b && pe_ch && out), "d2f_pressure_dp_f64: null pointer");
  if (n_pe == 0) return 0;
  hipLaunchKernelGGL(k_pressure_dp, dim3((n_pe + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, lam, pe_a,
                     pe_b, pe_ch, n_pe, h2, out);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_ipc_hvp_f64(const double *x, const double *v, const int32_t *pairs, int32_t n_pairs,
                               const int32_t *ea, const int32_t *eb, const double *ca, const double *cb,
                               double dhat, double kappa, double h2, double *y, void *stream) {
  D2F_CHECK(n_pairs >= 0 && dhat > 0.0, "d2f_ipc_hvp_f64: bad argument");
  if (n_pairs == 0) return 0;
  D2F_CHECK(x && v && pairs && ea && eb && ca && cb && y, "d2f_ipc_hvp_f64: null pointer");
  hipLaunchKernelGGL(k_ipc_hvp, dim3((n_pairs + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, v, pairs,
                     n_pairs, ea, eb, ca, cb, dhat, kappa, h2, y);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_ipc_grad_diag_f64(const double *x, const int32_t *pairs, int32_t n_pairs, const int32_t *ea,
                                     const int32_t *eb, const double *ca, const double *cb, double dhat,
                                     double kappa, double h2, double *g, double *dg, void *stream) {
  D2F_CHECK(n_pairs >= 0 && dhat > 0.0, "d2f_ipc_grad_diag_f64: bad argument");
  if (n_pairs == 0) return 0;
  D2F_CHECK(x && pairs && ea && eb && ca && cb && g && dg, "d2f_ipc_grad_diag_f64: null pointer");
  hipLaunchKernelGGL(k_ipc_grad_diag, dim3((n_pairs + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, pairs,
                     n_pairs, ea, eb, ca, cb, dhat, kappa, h2, g, dg);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_affine_expand_f64(const double *vr, const int32_t *slave_of, const int32_t *an_grp,
                                     const double *an_w, const int32_t *slot, int32_t N, double *vf,
                                     void *stream) {
  D2F_CHECK(N >= 0, "d2f_affine_expand_f64: bad argument");
  if (N == 0) return 0;
  D2F_CHECK(vr && slave_of && an_grp && an_w && slot && vf && vr != vf, "d2f_affine_expand_f64: null or aliased pointer");
  hipLaunchKernelGGL(k_affine_expand, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, vr, slave_of,
                     an_grp, an_w, slot, N, vf);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_affine_reduce_f64(const double *yf, const double *vf, const double *mass,
                                     const int32_t *slave_of, const int32_t *an_node, const double *an_w,
                                     const int32_t *an_first, const int32_t *slot, int32_t N, int32_t n_groups,
                                     int32_t square, double *yr, void *stream) {
  D2F_CHECK(N >= 0 && n_groups >= 0, "d2f_affine_reduce_f64: bad argument");
  if (N == 0) return 0;
  D2F_CHECK(yf && slave_of && yr && yf != yr && (!vf || mass) && (n_groups == 0 || (an_node && an_w && an_first && slot)),
            "d2f_affine_reduce_f64: null or aliased pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_affine_reduce_nodes, dim3((N + 255) / 256), dim3(256), 0, st, yf, slave_of, N, yr);
  if (n_groups > 0)
    hipLaunchKernelGGL(k_affine_reduce_groups, dim3(n_groups), dim3(AR_TPB), 0, st, yf, vf, mass, an_node, an_w,
                       an_first, slot, square, yr);
  D2F_HIP(hipGetLastError());
  return 0;
}

static int minres_check(const d2f_minres *m, const char *who) {
  D2F_CHECK(m && m->B >= 0, "%s: bad argument", who);
  D2F_CHECK(m->B == 0 || (m->node_start && m->mass && m->fmask && m->dinv && m->b && m->x && m->v && m->r1 &&
                          m->r2 && m->w1 && m->w2 && m->Kv && m->scal && m->status),
            "%s: null pointer", who);
  return 0;
}

extern "C" int d2f_minres_begin_f64(const d2f_minres *m, void *stream) {
  if (int rc = minres_check(m, "d2f_minres_begin_f64")) return rc;
  if (m->B == 0) return 0;
  hipLaunchKernelGGL(k_minres_begin, dim3(m->B), dim3(MR_TPB), 0, (hipStream_t)stream, *m);
  D2F_HIP(hipGetLastError());
  return 0;
}

extern "C" int d2f_minres_step_f64(const d2f_minres *m, void *stream) {
  if (int rc = minres_check(m, "d2f_minres_step_f64")) return rc;
  if (m->B == 0) return 0;
  hipLaunchKernelGGL(k_minres_step, dim3(m->B), dim3(MR_TPB), 0, (hipStream_t)stream, *m);
  D2F_HIP(hipGetLastError());
  return 0;
}
