// "node" contact between bodies: one visit of a boundary node to a partner (reference
// implicit/kernels.py:755-869, implicit/body_contact.py:37-167, implicit/adjoint.py:115-143).
//
// Two boundary nodes of one scene closer than r, and not excluded against each other, repel with
// f_n = k (r - d) along n = (x_a - x_b) / d, d = sqrt(|x_a - x_b|^2 + 1e-12). Every pair is
// visited from both of its nodes. A visit adds its own node's gradient and diagonal, and half
// of the pair's p^T H p and energy. Friction on the pair is the colliders' smoothed Coulomb law
// on the relative slip over the step, u = (x_a - xn_a) - (x_b - xn_b).
#pragma once
#include "common.h"

namespace d2f {

constexpr double NODE_REG = 1.0e-12;

struct NodeVisit {
  V2<double> n;  // from the partner to the visiting node, unit up to the regulariser
  double d, rd;  // regularised distance, r - d
};

__device__ __forceinline__ bool node_visit(V2<double> xa, V2<double> xb, double r, NodeVisit &v) {
  const V2<double> diff = xa - xb;
  v.d = sqrt(dot(diff, diff) + NODE_REG);
  if (!(v.d < r)) return false;
  v.n = {diff.x / v.d, diff.y / v.d};
  v.rd = r - v.d;
  return true;
}

// penalty gradient and diagonal of the visiting node (reference implicit/kernels.py:796-806)
__device__ __forceinline__ void node_grad_diag(const NodeVisit &v, double k, int clamp, V2<double> &g,
                                               double &hx, double &hy) {
  g = g - (k * v.rd) * v.n;
  const double tang = k * v.rd / v.d;
  double ax = k * v.n.x * v.n.x - tang * (1.0 - v.n.x * v.n.x);
  double ay = k * v.n.y * v.n.y - tang * (1.0 - v.n.y * v.n.y);
  if (clamp) {
    ax = fmax(ax, 0.0);
    ay = fmax(ay, 0.0);
  }
  hx += ax;
  hy += ay;
}

// half of the pair's p^T H p along the relative direction dp = p_a - p_b (kernels.py:834-843)
__device__ __forceinline__ double node_pHp(const NodeVisit &v, double k, int clamp, V2<double> dp) {
  const double pn = dot(dp, v.n);
  const V2<double> pt = dp - pn * v.n;
  double val = k * pn * pn - k * v.rd / v.d * dot(pt, pt);
  if (clamp) val = fmax(val, 0.0);
  return 0.5 * val;
}

// half of the pair's energy 1/2 k (r - d)^2 (kernels.py:864-867)
__device__ __forceinline__ double node_energy(const NodeVisit &v, double k) {
  return 0.5 * (0.5 * k * v.rd * v.rd);
}

// own row of H v, unclamped (reference implicit/adjoint.py:136-142)
__device__ __forceinline__ V2<double> node_hvp(const NodeVisit &v, double k, V2<double> dv) {
  const double vn = dot(dv, v.n);
  const V2<double> vt = dv - vn * v.n;
  return (k * vn) * v.n - (k * v.rd / v.d) * vt;
}

// Coefficient of the pair (reference core/math.py:133-144): mu, or with per-body values the
// harmonic mean of the two nodes' bodies' coefficients (0 when both are 0).
// (Any pointer type: a plain pointer or one in the global address space.)
template <typename Ptr>
__device__ __forceinline__ double node_pair_mu(double mu, int per_body, Ptr mu_node, int i, int j) {
  if (!per_body || !mu_node) return mu;
  const double mi = mu_node[i], mj = mu_node[j], sm = mi + mj;
  return sm > 0.0 ? 2.0 * mi * mj / sm : 0.0;
}

// friction gradient and tangential stiffness of the visiting node (reference
// implicit/body_contact.py:63-77); du = u_a - u_b
__device__ __forceinline__ void node_fric_grad_diag(const NodeVisit &v, double k, double mu, double epsv,
                                                    double inv_dt, V2<double> du, V2<double> &g, double &fx,
                                                    double &fy) {
  const double f_n = k * v.rd;
  const V2<double> vel = inv_dt * du;
  const V2<double> vt = vel - dot(vel, v.n) * v.n;
  const double den = fmax(len(vt), epsv);
  g = g + (mu * f_n / den) * vt;
  const double kt = mu * f_n * inv_dt / den;
  fx += kt * v.n.y * v.n.y;
  fy += kt * v.n.x * v.n.x;
}

// half of the pair's friction curvature along dp (reference implicit/body_contact.py:110-121)
__device__ __forceinline__ double node_fric_pHp(const NodeVisit &v, double k, double mu, double epsv,
                                                double inv_dt, V2<double> du, V2<double> dp) {
  const double f_n = k * v.rd;
  const V2<double> vel = inv_dt * du;
  const V2<double> vt = vel - dot(vel, v.n) * v.n;
  const double kt = mu * f_n * inv_dt / fmax(len(vt), epsv);
  const double pt = dp.y * v.n.x - dp.x * v.n.y;
  return 0.5 * kt * pt * pt;
}

// half of the pair's friction potential mu f_n f0(|u_t|) (reference
// implicit/body_contact.py:154-165)
__device__ __forceinline__ double node_fric_energy(const NodeVisit &v, double k, double mu, double epsv,
                                                   double inv_dt, V2<double> du) {
  const double f_n = k * v.rd;
  const double a = epsv / inv_dt;
  const double ut = fabs(du.y * v.n.x - du.x * v.n.y);
  const double f0 = ut < a ? ut * ut / (2.0 * a) : ut - 0.5 * a;
  return 0.5 * mu * f_n * f0;
}

// A pass of the "node" contact over every scene's boundary list, one workgroup per scene
// (defined in step_plain.hip). mode 0: g, dg += h2 gradient, diagonal (friction too when fric);
// 1: acc[env] += h2 p^T H p along pdir; 2: acc[env] += h2 energy; 3: g += h2 H pdir (penalty,
// own rows, unclamped).
struct NodeArgs {
  const double *x, *xn, *pdir;
  const int32_t *bnd, *bnd_start;
  int32_t B, max_nb;
  const uint32_t *excl_bits;     // bit i * n_be + j of scene e's words: nodes i, j excluded
  const int64_t *excl_bits_off;  // B + 1 word offsets
  const uint8_t *freev;      // friction's free mask per node (NULL: every node free)
  const double *mu_node;     // per-node coefficient when per_body
  double k, r, mu, epsv, inv_dt, h2;
  int32_t clamp, fric, per_body;
  double *g, *dg, *acc;
};
int launch_node_env(int mode, const NodeArgs &a, hipStream_t st);

}  // namespace d2f
