// cpu_step.cpp — a serial C++ restatement of the oracle's implicit step (oracle/pncg.py,
// oracle/terms.py, oracle/ipc.py, oracle/bodies.py; reference implicit/solver.py:252-451 and the
// kernels it launches), TEST AND BASELINE INFRASTRUCTURE ONLY: the product never links it. It
// exists to time the reference algorithm on host cores as compiled code (the bench's CPU
// baseline) and is checked against the numpy oracle (tests/test_oracle_cpu.py).
//
// Scope: what the contact-heavy configuration uses — Neo-Hookean elasticity, half-space and
// disc colliders with lagged friction, IPC node-edge contact between bodies with lagged friction
// and vertex corrections, the kinematic sweep, affine (rigid) bodies, an external force field,
// the Dai-Kou direction with the restart guard, the step caps and the per-scene convergence
// test. Environments are independent: each is stepped on its own (serially, or one per OpenMP
// thread), and within one the loops run in index order like the reference's CPU kernels.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double BIG = 1.0e30;
constexpr double LOG_EPS = 1.0e-6;
constexpr double BC_FLOOR = 0.25;

struct V {
  double x, y;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y}; }
inline V operator*(double s, V a) { return {s * a.x, s * a.y}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y; }
inline double cross(V a, V b) { return a.x * b.y - a.y * b.x; }
inline double norm(V a) { return std::sqrt(a.x * a.x + a.y * a.y); }

struct M {
  double a, b, c, d;  // [[a, b], [c, d]]
};
inline M mul(M p, M q) { return {p.a * q.a + p.b * q.c, p.a * q.b + p.b * q.d, p.c * q.a + p.d * q.c, p.c * q.b + p.d * q.d}; }
inline double det(M m) { return m.a * m.d - m.b * m.c; }
inline double ddot(M p, M q) { return p.a * q.a + p.b * q.b + p.c * q.c + p.d * q.d; }
inline M tr(M m) { return {m.a, m.c, m.b, m.d}; }
inline M inv(M m) {
  const double D = det(m);
  return {m.d / D, -m.b / D, -m.c / D, m.a / D};
}
inline M scale(double s, M m) { return {s * m.a, s * m.b, s * m.c, s * m.d}; }
inline M add(M p, M q) { return {p.a + q.a, p.b + q.b, p.c + q.c, p.d + q.d}; }

// Neo-Hookean (oracle/terms.py nh_P, nh_d2, nh_psi; reference core/math.py:76-94)
inline double logJ(M F) { return std::log(std::max(det(F), LOG_EPS)); }
inline M nh_P(M F, double mu, double lam) {
  const M G = tr(inv(F));
  return add(scale(mu, F), scale(lam * logJ(F) - mu, G));
}
inline M nh_d2(M F, M dF, double mu, double lam) {
  const M G = tr(inv(F));
  const double c = lam * logJ(F) - mu;
  return add(add(scale(mu, dF), scale(lam * ddot(G, dF), G)), scale(-c, mul(mul(G, tr(dF)), G)));
}

// barrier b(d) = -(d - dhat)^2 log(d / dhat) (oracle/terms.py; reference core/math.py:113-129)
inline double bp(double d, double dh) {
  const double u = d - dh;
  return -2.0 * u * std::log(d / dh) - u * u / d;
}
inline double bpp(double d, double dh) {
  const double u = d - dh;
  return -2.0 * std::log(d / dh) - 4.0 * u / d + u * u / (d * d);
}

}  // namespace

extern "C" {

// Flat scene, grouped by environment (every *_start has B + 1 entries). Node vectors are
// interleaved (x, y); Bm is row-major. Indices are int64. ela / elb and excl_idx are
// environment-local boundary indices. Colliders: 8 doubles each (kind, cx, cy, rx, ry, rref,
// r, mu).
struct d2f_cpu_scene {
  int64_t B, N, T, n_b, n_e, n_bc, n_ag, n_col;
  double *x, *v, *xn, *xhat, *grad, *grad_prev, *p, *diag, *bc_q;  // 2N
  const double *mass, *free_, *fext;                                // N, 2N, 2N (fext may be NULL)
  const int64_t *node_start;
  const int64_t *tri;                                               // 3T
  const double *Bm, *area, *mu, *lam;                               // 4T, T, T, T
  const int64_t *tri_start;
  const int64_t *bnd, *bnd_start;
  const int64_t *ea, *eb, *ela, *elb, *e_start;
  const double *ca, *cb;
  const int64_t *excl_ptr;                                          // n_b + 1
  const int32_t *excl_idx;
  const double *mu_node;                                            // N or NULL
  const int64_t *bc_idx, *bc_start;                                 // n_bc
  const double *bc_target;                                          // 2 n_bc
  const int64_t *an_node, *an_start, *ag_start;                     // slaved nodes, groups
  const double *an_w, *ag_Mu, *ag_La;                               // 3 per slaved node
  double *ag_q;                                                     // 6 per group (state)
  const double *col;
  int32_t *iters, *active;                                          // B
};

struct d2f_cpu_params {
  double dt, drag, dhat, kappa, epsv, eps, tol_x, near_factor, alpha_max, area_shrink, ccd_eta, cg_guard;
  double ipc_mu, ipc_dhat, ipc_kappa, ipc_reach, ipc_gap;
  int32_t iter_max, conv_check, friction, ipc_on, ipc_per_body, pad_;
};

}  // extern "C"

namespace {

using S = d2f_cpu_scene;
using P = d2f_cpu_params;

inline V ld(const double *a, int64_t i) { return {a[2 * i], a[2 * i + 1]}; }
inline void st(double *a, int64_t i, V v) {
  a[2 * i] = v.x;
  a[2 * i + 1] = v.y;
}
inline void addv(double *a, int64_t i, V v) {
  a[2 * i] += v.x;
  a[2 * i + 1] += v.y;
}

inline double seg_dist(V p, V a, V b) {
  const V ab = b - a;
  const double t = std::min(std::max(dot(p - a, ab) / std::max(dot(ab, ab), 1.0e-30), 0.0), 1.0);
  return norm(p - a - t * ab);
}

// additive CCD of one node against one edge (oracle/ipc.py accd; reference
// implicit/ipc_contact.py:103-127)
double accd(V p, V a, V b, V dp, V da, V db, double g, double t_c) {
  const V m = (1.0 / 3.0) * (dp + da + db);
  dp = dp - m;
  da = da - m;
  db = db - m;
  const double lp = norm(dp) + std::max(norm(da), norm(db));
  if (!(lp > 1.0e-30)) return t_c;
  double t = 0.0, tl = (seg_dist(p, a, b) - g) / lp;
  for (int k = 0; k < 32; ++k) {
    const double tn = t + tl;
    if (tn >= t_c) return t_c;
    const double dn = seg_dist(p + tn * dp, a + tn * da, b + tn * db);
    if (k > 0 && dn < g) return t;
    t = tn;
    tl = 0.9 * dn / lp;
  }
  return t;
}

struct Env {
  const S &s;
  const P &q;
  int64_t e, n0, n1, t0, t1, b0, b1, k0, k1, c0, c1, g0, g1;
  double h2, inv_dt;
  // per-environment scalars (oracle/pncg.py OracleSolver)
  double gPg = 0, g_p = 0, g_Py = 0, y_p = 0, y_Py = 0, beta = 0, gTp = 0, pHp = 0, alpha = 0, p_inf = 0,
         clr_min = BIG, dE = 0, dE0 = 0;
  int active = 1, iters = 0;
  // IPC sweep state (oracle/ipc.py)
  double q_inf = 0, sweep_beta = 1, clr_ipc = BIG;
  bool pending = false;
  // affine groups: q (in scene), pq, gq, gq_prev, zq (6 each), Hx, Hy (9 each)
  std::vector<double> pq, gq, gq_prev, zq, Hx, Hy;

  Env(const S &s_, const P &q_, int64_t e_) : s(s_), q(q_), e(e_) {
    n0 = s.node_start[e];
    n1 = s.node_start[e + 1];
    t0 = s.tri_start[e];
    t1 = s.tri_start[e + 1];
    b0 = s.bnd_start[e];
    b1 = s.bnd_start[e + 1];
    k0 = s.n_e ? s.e_start[e] : 0;
    k1 = s.n_e ? s.e_start[e + 1] : 0;
    c0 = s.n_bc ? s.bc_start[e] : 0;
    c1 = s.n_bc ? s.bc_start[e + 1] : 0;
    g0 = s.n_ag ? s.ag_start[e] : 0;
    g1 = s.n_ag ? s.ag_start[e + 1] : 0;
    h2 = q.dt * q.dt;
    inv_dt = 1.0 / q.dt;
    const size_t ng = (size_t)(g1 - g0);
    pq.assign(6 * ng, 0.0);
    gq.assign(6 * ng, 0.0);
    gq_prev.assign(6 * ng, 0.0);
    zq.assign(6 * ng, 0.0);
    Hx.assign(9 * ng, 0.0);
    Hy.assign(9 * ng, 0.0);
  }

  V fr(int64_t i) const { return ld(s.free_, i); }

  // ---------------------------------------------------------------- elasticity
  M Fof(int64_t t, V p0, V p1, V p2) const {
    const M Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
    const double *b = s.Bm + 4 * t;
    return mul(Ds, M{b[0], b[1], b[2], b[3]});
  }
  void elastic_grad_diag() {
    for (int64_t t = t0; t < t1; ++t) {
      const int64_t i[3] = {s.tri[3 * t], s.tri[3 * t + 1], s.tri[3 * t + 2]};
      if (s.mu[t] == 0.0 && s.lam[t] == 0.0) continue;      // rigid-region elements
      const M F = Fof(t, ld(s.x, i[0]), ld(s.x, i[1]), ld(s.x, i[2]));
      const double *bb = s.Bm + 4 * t;
      const M B = {bb[0], bb[1], bb[2], bb[3]};
      const M H = scale(s.area[t], mul(nh_P(F, s.mu[t], s.lam[t]), tr(B)));
      const V g1 = {H.a, H.c}, g2 = {H.b, H.d};
      addv(s.grad, i[0], h2 * V{-(g1.x + g2.x), -(g1.y + g2.y)});
      addv(s.grad, i[1], h2 * g1);
      addv(s.grad, i[2], h2 * g2);
      const double ab[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
      for (int n = 0; n < 3; ++n)
        for (int c = 0; c < 2; ++c) {
          M dDs = {0, 0, 0, 0};
          if (c == 0) {
            dDs.a = ab[n][0];
            dDs.b = ab[n][1];
          } else {
            dDs.c = ab[n][0];
            dDs.d = ab[n][1];
          }
          const M dF = mul(dDs, B);
          const double v = std::max(s.area[t] * ddot(dF, nh_d2(F, dF, s.mu[t], s.lam[t])), 0.0);
          s.diag[2 * i[n] + c] += h2 * v;
        }
    }
  }
  void elastic_pHp() {
    for (int64_t t = t0; t < t1; ++t) {
      if (s.mu[t] == 0.0 && s.lam[t] == 0.0) continue;
      const int64_t i[3] = {s.tri[3 * t], s.tri[3 * t + 1], s.tri[3 * t + 2]};
      const M F = Fof(t, ld(s.x, i[0]), ld(s.x, i[1]), ld(s.x, i[2]));
      const M dF = Fof(t, ld(s.p, i[0]), ld(s.p, i[1]), ld(s.p, i[2]));
      pHp += h2 * std::max(s.area[t] * ddot(dF, nh_d2(F, dF, s.mu[t], s.lam[t])), 0.0);
    }
  }
  void inversion_alpha() {
    for (int64_t t = t0; t < t1; ++t) {
      const V x0 = ld(s.x, s.tri[3 * t]), x1 = ld(s.x, s.tri[3 * t + 1]), x2 = ld(s.x, s.tri[3 * t + 2]);
      const V p0 = ld(s.p, s.tri[3 * t]), p1 = ld(s.p, s.tri[3 * t + 1]), p2 = ld(s.p, s.tri[3 * t + 2]);
      const V e1 = x1 - x0, e2 = x2 - x0, q1 = p1 - p0, q2 = p2 - p0;
      const double a0 = cross(e1, e2), a1 = cross(e1, q2) + cross(q1, e2), a2 = cross(q1, q2);
      const double cc = a0 - q.area_shrink * a0;
      double sres = BIG;
      if (std::fabs(a2) < 1.0e-24) {
        if (std::fabs(a1) > 1.0e-24 && -cc / a1 > 0.0) sres = -cc / a1;
      } else {
        const double disc = a1 * a1 - 4.0 * a2 * cc;
        if (disc >= 0.0) {
          const double sq = std::sqrt(disc);
          const double r1 = (-a1 - sq) / (2.0 * a2), r2 = (-a1 + sq) / (2.0 * a2);
          if (r1 > 0.0) sres = std::min(sres, r1);
          if (r2 > 0.0) sres = std::min(sres, r2);
        }
      }
      if (sres < BIG) alpha = std::min(alpha, sres);
    }
  }

  // ---------------------------------------------------------------- colliders
  // level set of collider k at X: (phi, grad phi); kinds 0 (disc) and 4 (half-space)
  bool col_eval(int k, V X, double &d, V &g, double &dc) const {
    const double *c = s.col + 8 * k;
    const int kind = (int)c[0];
    if (kind == 4) {
      g = {c[3], c[4]};
      d = dot(X - V{c[1], c[2]}, g);
      dc = 0.0;
      return true;
    }
    const V diff = X - V{c[1], c[2]};
    dc = norm(diff);
    d = dc - c[6];
    g = (1.0 / dc) * diff;
    return true;
  }
  void colliders_grad_diag() {
    for (int k = 0; k < s.n_col; ++k) {
      const bool disc = (int)s.col[8 * k] == 0;
      for (int64_t b = b0; b < b1; ++b) {
        const int64_t n = s.bnd[b];
        double d, dc;
        V g;
        col_eval(k, ld(s.x, n), d, g, dc);
        if (!(d > 0.0 && d < q.dhat)) continue;
        const double p1 = bp(d, q.dhat), p2 = bpp(d, q.dhat);
        addv(s.grad, n, (h2 * q.kappa * p1) * g);
        V dg = {q.kappa * p2 * g.x * g.x, q.kappa * p2 * g.y * g.y};
        if (disc) dg = dg + V{q.kappa * (p1 / dc) * (1.0 - g.x * g.x), q.kappa * (p1 / dc) * (1.0 - g.y * g.y)};
        addv(s.diag, n, h2 * V{std::max(dg.x, 0.0), std::max(dg.y, 0.0)});
      }
    }
  }
  // lagged friction against collider k at boundary node n: (f_n, unit normal, tangential v)
  bool fric_frame(int k, int64_t n, double &f_n, V &nh, V &vt) const {
    double d, dc;
    V g;
    col_eval(k, ld(s.x, n), d, g, dc);
    if (!(d > 0.0 && d < q.dhat)) return false;
    const double gn = norm(g);
    nh = (1.0 / std::max(gn, 1.0e-30)) * g;
    f_n = q.kappa * std::fabs(bp(d, q.dhat)) * ((int)s.col[8 * k] == 0 ? 1.0 : gn);
    const V v = inv_dt * (ld(s.x, n) - ld(s.xn, n));
    vt = v - dot(v, nh) * nh;
    return true;
  }
  bool free_gate(int64_t n) const {
    const V f = fr(n);
    return !(f.x <= 0.5 && f.y <= 0.5);
  }
  void friction_grad_diag() {
    for (int k = 0; k < s.n_col; ++k) {
      const double mu = s.col[8 * k + 7];
      for (int64_t b = b0; b < b1; ++b) {
        const int64_t n = s.bnd[b];
        double f_n;
        V nh, vt;
        if (!free_gate(n) || !fric_frame(k, n, f_n, nh, vt)) continue;
        const double vtm = norm(vt), den = std::max(vtm, q.epsv);
        const V f = fr(n);
        const V tdir = (1.0 / den) * vt;
        addv(s.grad, n, (h2 * mu * f_n) * V{tdir.x * f.x, tdir.y * f.y});
        const double kt = h2 * mu * f_n * inv_dt / den;
        const V th = {-nh.y, nh.x};
        addv(s.diag, n, V{kt * th.x * th.x * f.x, kt * th.y * th.y * f.y});
      }
    }
  }
  void colliders_pHp() {
    for (int k = 0; k < s.n_col; ++k) {
      const bool disc = (int)s.col[8 * k] == 0;
      for (int64_t b = b0; b < b1; ++b) {
        const int64_t n = s.bnd[b];
        double d, dc;
        V g;
        col_eval(k, ld(s.x, n), d, g, dc);
        clr_min = std::min(clr_min, d);
        if (!(d > 0.0 && d < q.dhat)) continue;
        const V pp = ld(s.p, n);
        const double pn = dot(pp, g);
        double val = q.kappa * bpp(d, q.dhat) * pn * pn;
        if (disc) {
          const V pt = pp - pn * g;
          val += q.kappa * (bp(d, q.dhat) / dc) * dot(pt, pt);
        }
        pHp += h2 * std::max(val, 0.0);
      }
    }
  }
  void friction_pHp() {
    for (int k = 0; k < s.n_col; ++k) {
      const double mu = s.col[8 * k + 7];
      for (int64_t b = b0; b < b1; ++b) {
        const int64_t n = s.bnd[b];
        double f_n;
        V nh, vt;
        if (!free_gate(n) || !fric_frame(k, n, f_n, nh, vt)) continue;
        const double kt = h2 * mu * f_n * inv_dt / std::max(norm(vt), q.epsv);
        const V f = fr(n), pp = ld(s.p, n);
        const double pt = pp.x * f.x * (-nh.y) + pp.y * f.y * nh.x;
        pHp += kt * pt * pt;
      }
    }
  }
  void colliders_ccd() {
    for (int k = 0; k < s.n_col; ++k) {
      const double *c = s.col + 8 * k;
      for (int64_t b = b0; b < b1; ++b) {
        const int64_t n = s.bnd[b];
        const V X = ld(s.x, n), Pp = ld(s.p, n);
        if ((int)c[0] == 0) {
          const V o = X - V{c[1], c[2]};
          const double a = dot(Pp, Pp), bq = 2.0 * dot(o, Pp), cq = dot(o, o) - c[6] * c[6];
          const double disc = bq * bq - 4.0 * a * cq;
          if (!(a >= 1.0e-24 && disc >= 0.0)) continue;
          const double sq = std::sqrt(disc);
          const double r1 = (-bq - sq) / (2.0 * a), r2 = (-bq + sq) / (2.0 * a);
          double sres = BIG;
          if (r1 > 0.0) sres = std::min(sres, r1);
          if (r2 > 0.0) sres = std::min(sres, r2);
          if (sres < BIG) alpha = std::min(alpha, q.ccd_eta * sres);
        } else {
          double d, dc;
          V g;
          col_eval(k, X, d, g, dc);
          const double gp = dot(g, Pp);
          if (!(gp >= 0.0 || d <= 0.0)) alpha = std::min(alpha, q.ccd_eta * d / (-gp));
        }
      }
    }
  }

  // ---------------------------------------------------------------- IPC between bodies
  bool excluded(int64_t jl, int64_t other_local) const {
    const int64_t r0 = s.excl_ptr[b0 + jl], r1 = s.excl_ptr[b0 + jl + 1];
    return std::binary_search(s.excl_idx + r0, s.excl_idx + r1, (int32_t)other_local);
  }
  double pair_mu(int64_t pn, int64_t a) const {
    if (!q.ipc_per_body) return q.ipc_mu;
    const double mi = s.mu_node[pn], mj = s.mu_node[a], sm = mi + mj;
    return sm > 0.0 ? 2.0 * mi * mj / sm : 0.0;
  }
  // one barrier-and-friction term (oracle/ipc.py _pair; reference implicit/ipc_contact.py:131-171)
  void ipc_term(int mode, double coef, int64_t ip, int64_t ia, int64_t ib, double wa, double wb, double d, V n,
                double ddd, double mu, const double *pdir) {
    const double b1 = bp(d, q.ipc_dhat);
    const double f_n = -q.ipc_kappa * b1;
    const V tn = {-n.y, n.x};
    const V u = (ld(s.x, ip) - ld(s.xn, ip)) - wa * (ld(s.x, ia) - ld(s.xn, ia)) - wb * (ld(s.x, ib) - ld(s.xn, ib));
    const double ut = dot(u, tn), vt = std::fabs(ut) * inv_dt;
    const double kt = mu * f_n * inv_dt / std::max(vt, q.epsv);
    if (mode == 0) {
      const double b2 = bpp(d, q.ipc_dhat);
      const V g = h2 * ((-f_n) * n + (mu * f_n * (ut * inv_dt / std::max(vt, q.epsv))) * tn);
      const V hd = {h2 * (q.ipc_kappa * b2 * n.x * n.x + kt * tn.x * tn.x),
                    h2 * (q.ipc_kappa * b2 * n.y * n.y + kt * tn.y * tn.y)};
      const V fp = fr(ip), fa = fr(ia), fb = fr(ib);
      addv(s.grad, ip, coef * V{g.x * fp.x, g.y * fp.y});
      addv(s.diag, ip, coef * hd);
      addv(s.grad, ia, (-coef * wa) * V{g.x * fa.x, g.y * fa.y});
      addv(s.diag, ia, (coef * wa * wa) * hd);
      addv(s.grad, ib, (-coef * wb) * V{g.x * fb.x, g.y * fb.y});
      addv(s.diag, ib, (coef * wb * wb) * hd);
    } else {
      const V dq = ld(pdir, ip) - wa * ld(pdir, ia) - wb * ld(pdir, ib);
      const double dn = dot(dq, n), dt_ = dot(dq, tn);
      const double val = std::max(q.ipc_kappa * (bpp(d, q.ipc_dhat) * dn * dn + b1 * ddd), 0.0);
      pHp += coef * h2 * (val + kt * dt_ * dt_);
    }
  }
  // a pass over every (edge, boundary node) pair of the scene within reach (oracle/ipc.py _run;
  // reference implicit/ipc_contact.py:175-296). mode 0 gradient/diagonal, 1 curvature with
  // clearance and step limit, 3 the sweep's step limit into alpha_out.
  void ipc_pass(int mode, const double *pdir, double mu, double floor_, double t_c, double &alpha_out) {
    const double dhat = q.ipc_dhat, dlo = 1.0e-9 * dhat, onset = 0.02 * dhat;
    for (int64_t k = k0; k < k1; ++k) {
      const int64_t a = s.ea[k], b = s.eb[k], la = s.ela[k], lb = s.elb[k];
      const V xa = ld(s.x, a), xb = ld(s.x, b), ab = xb - xa;
      const double L2 = std::max(dot(ab, ab), 1.0e-30), L = std::sqrt(L2);
      for (int64_t jl = 0; jl < b1 - b0; ++jl) {
        const int64_t pn = s.bnd[b0 + jl];
        if (pn == a || pn == b) continue;
        const V xp = ld(s.x, pn), w = xp - xa;
        const double tt = dot(w, ab) / L2;
        V dvec = w - tt * ab;
        if (tt <= 0.0) dvec = w;
        if (tt >= 1.0) dvec = xp - xb;
        const double d = norm(dvec);
        if (!(d < q.ipc_reach)) continue;
        if (excluded(jl, la) || excluded(jl, lb)) continue;
        if (mode == 1 || mode == 3) {
          clr_ipc = std::min(clr_ipc, d);
          double keep = std::max(q.ipc_gap * d, dhat - onset);
          if (d < dhat) keep = d - std::min(2.0 * (dhat - d) + onset, 0.5 * d);
          if (mode == 3) keep = std::min(d, std::max(q.ipc_gap * d, floor_));
          alpha_out = std::min(alpha_out, accd(xp, xa, xb, ld(pdir, pn), ld(pdir, a), ld(pdir, b), keep, t_c));
        }
        if (mode == 3) continue;
        const double mp = pair_mu(pn, a);
        if (tt > 0.0 && tt < 1.0 && d < dhat) {
          const double di = std::max(d, dlo);
          double ddd = 0.0;
          if (mode == 1) {
            const V de = ld(pdir, b) - ld(pdir, a), dw = ld(pdir, pn) - ld(pdir, a);
            const double c0 = cross(ab, w), c1 = cross(de, w) + cross(ab, dw), c2 = 2.0 * cross(de, dw);
            const double L1 = dot(ab, de) / L, Lq = (dot(de, de) - L1 * L1) / L;
            const double s0 = c0 / L, s1 = (c1 - s0 * L1) / L;
            ddd = (c2 - 2.0 * s1 * L1 - s0 * Lq) / L;
            if (c0 < 0.0) ddd = -ddd;
          }
          ipc_term(mode, 1.0, pn, a, b, 1.0 - tt, tt, di, (1.0 / di) * dvec, ddd, mp, pdir);
        }
        for (int end = 0; end < 2; ++end) {
          const double coef = (end == 0 ? -s.ca[k] : -s.cb[k]) + ((end == 0 ? tt <= 0.0 : tt >= 1.0) ? 1.0 : 0.0);
          const V wv = end == 0 ? w : xp - xb;
          const double dd = norm(wv);
          if (coef == 0.0 || !(dd < dhat)) continue;
          const double de_ = std::max(dd, dlo);
          const V nn = (1.0 / de_) * wv;
          double ddd = 0.0;
          if (mode == 1) {
            const V dw = ld(pdir, pn) - ld(pdir, end == 0 ? a : b);
            const double dn = dot(dw, nn);
            ddd = (dot(dw, dw) - dn * dn) / de_;
          }
          ipc_term(mode, coef, pn, a, b, end == 0 ? 1.0 : 0.0, end == 0 ? 0.0 : 1.0, de_, nn, ddd, mp, pdir);
        }
      }
    }
  }
  // kinematic sweep (oracle/ipc.py _sweep; reference implicit/ipc_contact.py:308-331, 456-463)
  void sweep() {
    q_inf = 0.0;
    for (int64_t c = c0; c < c1; ++c) {
      const int64_t i = s.bc_idx[c];
      const V dq = ld(s.bc_target, c) - ld(s.x, i);
      st(s.bc_q, i, dq);
      q_inf = std::max(q_inf, norm(dq));
    }
    sweep_beta = 1.0;
    clr_ipc = BIG;
    ipc_pass(3, s.bc_q, 0.0, BC_FLOOR * q.ipc_dhat, 1.0, sweep_beta);
    if (q_inf > 0.0) sweep_beta = std::min(sweep_beta, 0.45 * q.ipc_reach / q_inf);
    for (int64_t c = c0; c < c1; ++c) {
      const int64_t i = s.bc_idx[c];
      const V tg = ld(s.bc_target, c);
      const V xi = sweep_beta >= 1.0 ? tg : ld(s.x, i) + sweep_beta * (tg - ld(s.x, i));
      st(s.x, i, xi);
      st(s.xhat, i, xi);
    }
  }

  // ---------------------------------------------------------------- affine bodies
  // (oracle/bodies.py AffineBodies; reference implicit/affine.py:61-269)
  double *qg(int64_t g) const { return s.ag_q + 6 * g; }
  void expand(int64_t g, const double *v6, double *out) const {
    for (int64_t j = s.an_start[g]; j < s.an_start[g + 1]; ++j) {
      const double *w = s.an_w + 3 * j;
      st(out, s.an_node[j], V{w[0] * v6[0] + w[1] * v6[1] + w[2] * v6[2], w[0] * v6[3] + w[1] * v6[4] + w[2] * v6[5]});
    }
  }
  void gather() {
    for (int64_t g = g0; g < g1; ++g) {
      const size_t o = (size_t)(g - g0);
      for (int k = 0; k < 6; ++k) {
        gq_prev[6 * o + k] = gq[6 * o + k];
        gq[6 * o + k] = 0.0;
      }
      for (int k = 0; k < 9; ++k) Hx[9 * o + k] = Hy[9 * o + k] = 0.0;
      for (int64_t j = s.an_start[g]; j < s.an_start[g + 1]; ++j) {
        const int64_t n = s.an_node[j];
        const double *w = s.an_w + 3 * j;
        const V gg = ld(s.grad, n), dd = ld(s.diag, n);
        for (int r = 0; r < 3; ++r) {
          gq[6 * o + r] += gg.x * w[r];
          gq[6 * o + 3 + r] += gg.y * w[r];
          for (int c = 0; c < 3; ++c) {
            Hx[9 * o + 3 * r + c] += dd.x * w[r] * w[c];
            Hy[9 * o + 3 * r + c] += dd.y * w[r] * w[c];
          }
        }
        st(s.grad, n, V{0.0, 0.0});
      }
    }
  }
  static void chol6(const double H[36], double L[36]) {
    bool ok = true;
    std::memset(L, 0, 36 * sizeof(double));
    for (int j = 0; j < 6 && ok; ++j) {
      double sj = H[6 * j + j];
      for (int k = 0; k < j; ++k) sj -= L[6 * j + k] * L[6 * j + k];
      if (sj <= 1.0e-9 * std::fabs(H[6 * j + j]) || sj <= 0.0) {
        ok = false;
        break;
      }
      const double dj = std::sqrt(sj);
      L[6 * j + j] = dj;
      for (int i = j + 1; i < 6; ++i) {
        double v = H[6 * i + j];
        for (int k = 0; k < j; ++k) v -= L[6 * i + k] * L[6 * j + k];
        L[6 * i + j] = v / dj;
      }
    }
    if (!ok) {
      std::memset(L, 0, 36 * sizeof(double));
      for (int j = 0; j < 6; ++j) L[6 * j + j] = std::sqrt(std::max(H[6 * j + j], 1.0e-30));
    }
  }
  static void chol6_solve(const double L[36], const double b[6], double x[6]) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
      double v = b[i];
      for (int k = 0; k < i; ++k) v -= L[6 * i + k] * y[k];
      y[i] = v / L[6 * i + i];
    }
    for (int i = 5; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < 6; ++k) v -= L[6 * k + i] * x[k];
      x[i] = v / L[6 * i + i];
    }
  }
  static M A_of(const double *v) { return {v[1], v[2], v[4], v[5]}; }
  void reduce() {
    static const int AIDX[4] = {1, 2, 4, 5};
    for (int64_t g = g0; g < g1; ++g) {
      const size_t o = (size_t)(g - g0);
      const M A = A_of(qg(g));
      const double Mu = s.ag_Mu[g], La = s.ag_La[g];
      const M Pa = nh_P(A, Mu, La);
      double gg[6], y[6], z[6], zy[6], H[36], L[36];
      for (int k = 0; k < 6; ++k) gg[k] = gq[6 * o + k];
      gg[1] += h2 * Pa.a;
      gg[2] += h2 * Pa.b;
      gg[4] += h2 * Pa.c;
      gg[5] += h2 * Pa.d;
      std::memset(H, 0, sizeof(H));
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          H[6 * r + c] = Hx[9 * o + 3 * r + c];
          H[6 * (3 + r) + 3 + c] = Hy[9 * o + 3 * r + c];
        }
      for (int k = 0; k < 4; ++k) {
        M E = {0, 0, 0, 0};
        (&E.a)[k] = 1.0;
        const M col = nh_d2(A, E, Mu, La);
        for (int r = 0; r < 4; ++r) H[6 * AIDX[r] + AIDX[k]] += h2 * (&col.a)[r];
      }
      // the negative twist direction removed (reference implicit/affine.py:61-101)
      const double ta = A.a + A.d, tb = A.c - A.b, nn = std::sqrt(ta * ta + tb * tb);
      if (nn > 1.0e-12) {
        const double f = 0.7071067811865476 / nn;
        const double t[4] = {-tb * f, -ta * f, ta * f, -tb * f};
        const M Tm = {t[0], t[1], t[2], t[3]};
        const double lt = ddot(Tm, nh_d2(A, Tm, Mu, La));
        if (lt < 0.0)
          for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) H[6 * AIDX[r] + AIDX[c]] -= h2 * lt * t[r] * t[c];
      }
      chol6(H, L);
      for (int k = 0; k < 6; ++k) y[k] = gg[k] - gq_prev[6 * o + k];
      chol6_solve(L, gg, z);
      chol6_solve(L, y, zy);
      for (int k = 0; k < 6; ++k) {
        gPg += gg[k] * z[k];
        g_p += gg[k] * pq[6 * o + k];
        g_Py += gg[k] * zy[k];
        y_p += y[k] * pq[6 * o + k];
        y_Py += y[k] * zy[k];
        gq[6 * o + k] = gg[k];
        zq[6 * o + k] = z[k];
      }
    }
  }
  void affine_direction() {
    for (int64_t g = g0; g < g1; ++g) {
      const size_t o = (size_t)(g - g0);
      for (int k = 0; k < 6; ++k) pq[6 * o + k] = -zq[6 * o + k] + beta * pq[6 * o + k];
      expand(g, &pq[6 * o], s.p);
    }
  }
  void affine_advance() {
    if (!active) return;
    for (int64_t g = g0; g < g1; ++g) {
      const size_t o = (size_t)(g - g0);
      for (int k = 0; k < 6; ++k) qg(g)[k] += alpha * pq[6 * o + k];
      expand(g, qg(g), s.x);
    }
  }
  void affine_pHp() {
    for (int64_t g = g0; g < g1; ++g) {
      const size_t o = (size_t)(g - g0);
      const M dA = A_of(&pq[6 * o]);
      pHp += h2 * std::max(ddot(dA, nh_d2(A_of(qg(g)), dA, s.ag_Mu[g], s.ag_La[g])), 0.0);
    }
  }

  // ---------------------------------------------------------------- one iteration
  void grad_diag() {
    for (int64_t i = n0; i < n1; ++i) {
      st(s.grad_prev, i, ld(s.grad, i));
      const V f = fr(i), d = ld(s.x, i) - ld(s.xhat, i);
      const double m = s.mass[i];
      st(s.grad, i, V{f.x * m * d.x, f.y * m * d.y});
      st(s.diag, i, V{m, m});
    }
    elastic_grad_diag();
    colliders_grad_diag();
    if (q.friction) friction_grad_diag();
    if (s.fext)
      for (int64_t i = n0; i < n1; ++i) addv(s.grad, i, (-h2) * ld(s.fext, i));
    double unused = 0.0;
    if (q.ipc_on) ipc_pass(0, s.p, q.ipc_mu, 0.0, 0.0, unused);
    gather();
    for (int64_t i = n0; i < n1; ++i) {
      const V f = fr(i), g = ld(s.grad, i);
      st(s.grad, i, V{g.x * f.x, g.y * f.y});
    }
  }
  void direction(int it) {
    gPg = g_p = g_Py = y_p = y_Py = 0.0;
    for (int64_t i = n0; i < n1; ++i) {
      const V f = fr(i), dg = ld(s.diag, i), g = ld(s.grad, i), pp = ld(s.p, i);
      const double Px = f.x / std::max(dg.x, 1.0e-12), Py = f.y / std::max(dg.y, 1.0e-12);
      const V yy = g - ld(s.grad_prev, i);
      gPg += g.x * Px * g.x + g.y * Py * g.y;
      g_p += g.x * pp.x + g.y * pp.y;
      g_Py += g.x * Px * yy.x + g.y * Py * yy.y;
      y_p += yy.x * pp.x + yy.y * pp.y;
      y_Py += yy.x * Px * yy.x + yy.y * Py * yy.y;
    }
    reduce();
    double b = 0.0;
    if (it > 0 && std::fabs(y_p) > 1.0e-30) b = g_Py / y_p - y_Py * g_p / (y_p * y_p);
    const double t = -gPg + b * g_p;
    if (t >= 0.0) {
      beta = 0.0;
      gTp = -gPg;
    } else {
      beta = b;
      gTp = t;
    }
    if (q.cg_guard > 0.0 && it > 0) {
      const double hold = (q.ipc_on && pending) ? q_inf : 0.0;
      if (hold <= 0.0 && y_p < q.cg_guard * alpha * pHp) {
        beta = 0.0;
        gTp = -gPg;
      }
    }
    for (int64_t i = n0; i < n1; ++i) {
      const V f = fr(i), dg = ld(s.diag, i), g = ld(s.grad, i), pp = ld(s.p, i);
      const double px = -f.x / std::max(dg.x, 1.0e-12) * g.x + beta * pp.x;
      const double py = -f.y / std::max(dg.y, 1.0e-12) * g.y + beta * pp.y;
      st(s.p, i, V{px * f.x, py * f.y});
    }
    affine_direction();
  }
  void curvature() {
    pHp = 0.0;
    alpha = q.alpha_max;
    p_inf = 0.0;
    clr_min = BIG;
    for (int64_t i = n0; i < n1; ++i) {
      const V pp = ld(s.p, i);
      pHp += s.mass[i] * dot(pp, pp);
      p_inf = std::max(p_inf, norm(pp));
    }
    elastic_pHp();
    colliders_pHp();
    if (q.friction) friction_pHp();
    affine_pHp();
    if (q.ipc_on) {
      clr_ipc = BIG;
      const double t_c = alpha;
      ipc_pass(1, s.p, q.ipc_mu, 0.0, t_c, alpha);
      if (p_inf > 0.0) alpha = std::min(alpha, 0.45 * q.ipc_reach / p_inf);
    }
  }
  void caps(int it) {
    inversion_alpha();
    colliders_ccd();
    if (!active) return;
    double a = alpha;
    if (pHp > 0.0) a = std::min(a, -gTp / pHp);
    if (clr_min < q.near_factor * q.dhat && p_inf > 0.0) a = std::min(a, 0.5 * q.dhat / p_inf);
    if (a <= 0.0) {
      active = 0;
      alpha = 0.0;
      return;
    }
    alpha = a;
    if (q.tol_x > 0.0 && a * p_inf < q.tol_x) {
      active = 0;
      ++iters;
      return;
    }
    const double d = -a * gTp - 0.5 * a * a * pHp;
    dE = d;
    if (it == 0) dE0 = std::max(d, 1.0e-300);
    else if (d < q.eps * dE0) active = 0;
    ++iters;
  }
  void update() {
    if (active)
      for (int64_t i = n0; i < n1; ++i) st(s.x, i, ld(s.x, i) + alpha * ld(s.p, i));
    affine_advance();
  }

  // one implicit step (oracle/pncg.py step_targets / iterate; reference implicit/solver.py:252-451)
  void step() {
    for (int64_t i = n0; i < n1; ++i) {
      const V x = ld(s.x, i);
      st(s.xn, i, x);
      st(s.xhat, i, x + (q.dt * q.drag) * ld(s.v, i));
    }
    pending = false;
    if (!q.ipc_on) {
      for (int64_t c = c0; c < c1; ++c) {
        st(s.x, s.bc_idx[c], ld(s.bc_target, c));
        st(s.xhat, s.bc_idx[c], ld(s.bc_target, c));
      }
    } else if (c1 > c0) {
      sweep();
      pending = !(sweep_beta >= 1.0);
    }
    dE0 = 0.0;
    active = 1;
    iters = 0;
    for (int64_t i = n0; i < n1; ++i) {
      st(s.p, i, V{0.0, 0.0});
      st(s.grad, i, V{0.0, 0.0});
    }
    for (int it = 0; it < q.iter_max; ++it) {
      if (q.ipc_on && pending) {
        sweep();
        if (q_inf > 0.0) active = 1;
        if (it % 8 == 7) pending = !(sweep_beta >= 1.0);
      }
      // this scene can no longer change: converged, and no sweep can wake it
      if (!active && !(q.ipc_on && pending)) break;
      grad_diag();
      direction(it);
      curvature();
      caps(it);
      update();
    }
    for (int64_t i = n0; i < n1; ++i) st(s.v, i, inv_dt * (ld(s.x, i) - ld(s.xn, i)));
    s.iters[e] = iters;
    s.active[e] = active;
  }
};

}  // namespace

extern "C" {

// One implicit step of every environment; threads: OpenMP threads over environments (1:
// serial). Returns 0, or -1 on a bad argument.
int d2f_cpu_step(const d2f_cpu_scene *s, const d2f_cpu_params *p, int threads) {
  if (!s || !p || s->B < 0 || threads < 1) return -1;
  for (int64_t k = 0; k < s->n_col; ++k) {
    const int kind = (int)s->col[8 * k];
    if (kind != 0 && kind != 4) return -1;
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t e = 0; e < s->B; ++e) {
    Env env(*s, *p, e);
    env.step();
  }
  return 0;
}

int d2f_cpu_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
