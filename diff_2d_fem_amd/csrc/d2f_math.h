// d2f_math.h — device math of the implicit PNCG-IPC step, shared by the per-kernel and the
// scene-persistent paths. Element arithmetic is templated on its scalar type R (double, or
// float in mixed precision); every accumulator the solver reduces stays double.
//
// Conventions (reference core/math.py:1-30): Ds has the edge vectors as COLUMNS,
// Ds = [x1-x0 | x2-x0], F = Ds Bm with Bm = D_m^-1, 2x2 matrices row-major (a b; c d).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace d2f {

constexpr double BIG = 1.0e30;

template <typename R> struct V2 {
  R x, y;
};
template <typename R> struct M2 {
  R a, b, c, d;  // [[a, b], [c, d]]
};

template <typename R> __device__ __forceinline__ V2<R> v2(R x, R y) { return V2<R>{x, y}; }
template <typename R> __device__ __forceinline__ V2<R> operator+(V2<R> p, V2<R> q) { return {p.x + q.x, p.y + q.y}; }
template <typename R> __device__ __forceinline__ V2<R> operator-(V2<R> p, V2<R> q) { return {p.x - q.x, p.y - q.y}; }
template <typename R> __device__ __forceinline__ V2<R> operator*(R s, V2<R> p) { return {s * p.x, s * p.y}; }
template <typename R> __device__ __forceinline__ R dot(V2<R> p, V2<R> q) { return p.x * q.x + p.y * q.y; }
template <typename R> __device__ __forceinline__ R cross(V2<R> p, V2<R> q) { return p.x * q.y - p.y * q.x; }
template <typename R> __device__ __forceinline__ R len(V2<R> p) { return sqrt(p.x * p.x + p.y * p.y); }

__device__ __forceinline__ V2<double> ld2(const double *a, int i) {
  const double2 t = reinterpret_cast<const double2 *>(a)[i];
  return {t.x, t.y};
}
__device__ __forceinline__ void st2(double *a, int i, V2<double> v) {
  reinterpret_cast<double2 *>(a)[i] = make_double2(v.x, v.y);
}
__device__ __forceinline__ void atomic_add2(double *a, int i, V2<double> v) {
  atomicAdd(a + 2 * i, v.x);
  atomicAdd(a + 2 * i + 1, v.y);
}
// Bit i of a bit-packed set (32 per word, little-endian within the word).
__device__ __forceinline__ bool bit_of(const uint32_t *w, int64_t i) { return (w[i >> 5] >> (i & 31)) & 1u; }
// The same on pointers known to be in global memory (address space 1).
#ifndef D2F_GLOBAL
#define D2F_GLOBAL __attribute__((address_space(1)))
#endif
__device__ __forceinline__ bool bit_of(const D2F_GLOBAL uint32_t *w, int64_t i) { return (w[i >> 5] >> (i & 31)) & 1u; }
typedef double d2f_dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ V2<double> ld2(const D2F_GLOBAL double *a, int i) {
  const d2f_dvec2 t = reinterpret_cast<const D2F_GLOBAL d2f_dvec2 *>(a)[i];
  return {t.x, t.y};
}
__device__ __forceinline__ void st2(D2F_GLOBAL double *a, int i, V2<double> v) {
  d2f_dvec2 t;
  t.x = v.x;
  t.y = v.y;
  reinterpret_cast<D2F_GLOBAL d2f_dvec2 *>(a)[i] = t;
}
// float64 atomic add on a global pointer
__device__ __forceinline__ void atomic_add_g(D2F_GLOBAL double *a, double v) { atomicAdd((double *)a, v); }
// min / max of non-negative-or-any doubles by compare-and-swap (deterministic result).
__device__ __forceinline__ void atomic_min_d(double *a, double v) {
  unsigned long long *p = reinterpret_cast<unsigned long long *>(a);
  unsigned long long old = *p;
  while (__longlong_as_double(old) > v) {
    unsigned long long prev = atomicCAS(p, old, __double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}
__device__ __forceinline__ void atomic_max_d(double *a, double v) {
  unsigned long long *p = reinterpret_cast<unsigned long long *>(a);
  unsigned long long old = *p;
  while (__longlong_as_double(old) < v) {
    unsigned long long prev = atomicCAS(p, old, __double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

// ------------------------------------------------------------------------------- Neo-Hookean
// Psi = mu/2 (|F|^2 - 2) - mu log J + lam/2 log^2 J, log J floored at J >= 1e-6
// (reference core/math.py:15-22, 76-100).
template <typename R> struct NH {
  M2<R> F;   // deformation gradient
  M2<R> G;   // F^-T
  M2<R> D;   // F - F^-T (the gradient's mu term)
  R logJ;
  R lL;      // lam log J
  R c;       // lam log J - mu
};

template <typename R>
__device__ __forceinline__ NH<R> nh_eval(M2<R> F, R mu, R lam) {
  NH<R> h;
  h.F = F;
  const R J = F.a * F.d - F.b * F.c;
  const R iJ = R(1) / J;
  h.G = {F.d * iJ, -F.c * iJ, -F.b * iJ, F.a * iJ};
  h.D = {F.a - h.G.a, F.b - h.G.b, F.c - h.G.c, F.d - h.G.d};
  h.logJ = log(J > R(1.0e-6) ? J : R(1.0e-6));
  h.lL = lam * h.logJ;
  h.c = h.lL - mu;
  return h;
}

// The same from the displacement gradient H = F - I, for float arithmetic near the rest state.
// Formed from F directly, J, log J and F - F^-T are differences of O(1) terms, so a float
// keeps only about 6e-8 / strain of their relative accuracy. From H, with s = J - 1 =
// tr H + det H:
//   log J = log1p(s)
//   (F - F^-T)11 = (2 H11 + det H + s H11) / J,  (F - F^-T)22 = (2 H22 + det H + s H22) / J,
//   (F - F^-T)12 = H12 + H21 / J,                 (F - F^-T)21 = H21 + H12 / J,
// every term small with the strain, so the stress keeps the float's relative accuracy.
template <typename R>
__device__ __forceinline__ NH<R> nh_eval_strain(M2<R> H, R mu, R lam) {
  NH<R> h;
  h.F = {R(1) + H.a, H.b, H.c, R(1) + H.d};
  const R detH = H.a * H.d - H.b * H.c;
  const R sJ = (H.a + H.d) + detH;
  const R J = R(1) + sJ;
  const R iJ = R(1) / J;
  h.G = {h.F.d * iJ, -H.c * iJ, -H.b * iJ, h.F.a * iJ};
  h.D = {(R(2) * H.a + detH + sJ * H.a) * iJ, H.b + H.c * iJ, H.c + H.b * iJ,
         (R(2) * H.d + detH + sJ * H.d) * iJ};
  h.logJ = J > R(1.0e-6) ? log1p(sJ) : log(R(1.0e-6));
  h.lL = lam * h.logJ;
  h.c = h.lL - mu;
  return h;
}

// dF : d2Psi/dF2 [dF] = mu |dF|^2 + lam (G:dF)^2 - c (dF : G dF^T G)
template <typename R>
__device__ __forceinline__ R nh_quad(const NH<R> &h, M2<R> dF, R mu, R lam) {
  const M2<R> &G = h.G;
  const R GdF = G.a * dF.a + G.b * dF.b + G.c * dF.c + G.d * dF.d;
  // T = G dF^T G ; dF:T
  const M2<R> GdFt = {G.a * dF.a + G.b * dF.b, G.a * dF.c + G.b * dF.d,
                      G.c * dF.a + G.d * dF.b, G.c * dF.c + G.d * dF.d};
  const M2<R> T = {GdFt.a * G.a + GdFt.b * G.c, GdFt.a * G.b + GdFt.b * G.d,
                   GdFt.c * G.a + GdFt.d * G.c, GdFt.c * G.b + GdFt.d * G.d};
  const R dd = dF.a * dF.a + dF.b * dF.b + dF.c * dF.c + dF.d * dF.d;
  const R dT = dF.a * T.a + dF.b * T.b + dF.c * T.c + dF.d * T.d;
  return mu * dd + lam * GdF * GdF - h.c * dT;
}

template <typename R>
__device__ __forceinline__ M2<R> mat_mul(M2<R> A, M2<R> B) {
  return {A.a * B.a + A.b * B.c, A.a * B.b + A.b * B.d, A.c * B.a + A.d * B.c, A.c * B.b + A.d * B.d};
}

// Elastic gradient of one element: the three nodal blocks of A P Bm^T (node 0 = -(g1+g2)).
template <typename R>
__device__ __forceinline__ void nh_grad(const NH<R> &h, M2<R> B, R mu, R A, V2<R> &g0, V2<R> &g1,
                                        V2<R> &g2) {
  const M2<R> &F = h.F, &G = h.G;
  M2<R> P;
  if constexpr (sizeof(R) == 4) {
    // float: P = mu (F - F^-T) + lam log J F^-T, from the cancellation-free F - F^-T
    const R l = h.lL;
    P = {mu * h.D.a + l * G.a, mu * h.D.b + l * G.b, mu * h.D.c + l * G.c, mu * h.D.d + l * G.d};
  } else {
    P = {mu * F.a + h.c * G.a, mu * F.b + h.c * G.b, mu * F.c + h.c * G.c, mu * F.d + h.c * G.d};
  }
  // (P B^T)_ij = P_i . B_j (rows)
  g1 = {A * (P.a * B.a + P.b * B.b), A * (P.c * B.a + P.d * B.b)};
  g2 = {A * (P.a * B.c + P.b * B.d), A * (P.c * B.c + P.d * B.d)};
  g0 = {-(g1.x + g2.x), -(g1.y + g2.y)};
}

// The six Jacobi entries of one element. A unit move of node n, component k, gives
// dF = e_k r_n^T with r_0 = -(B_0 + B_1), r_1 = B_0, r_2 = B_1 (B_i rows of Bm), so
// dF : d2Psi[dF] = mu |r|^2 + (lam - c) (G_k . r)^2.
template <typename R>
__device__ __forceinline__ void nh_diag(const NH<R> &h, M2<R> B, R mu, R lam, R A, V2<R> &d0,
                                        V2<R> &d1, V2<R> &d2) {
  const R k = lam - h.c;
  const V2<R> r1 = {B.a, B.b}, r2 = {B.c, B.d};
  const V2<R> r0 = {-(B.a + B.c), -(B.b + B.d)};
  const V2<R> Gx = {h.G.a, h.G.b}, Gy = {h.G.c, h.G.d};
  auto one = [&](V2<R> r) -> V2<R> {
    const R rr = mu * dot(r, r);
    const R sx = dot(Gx, r), sy = dot(Gy, r);
    return {A * (rr + k * sx * sx), A * (rr + k * sy * sy)};
  };
  d0 = one(r0);
  d1 = one(r1);
  d2 = one(r2);
}

template <typename R>
__device__ __forceinline__ R nh_psi(const NH<R> &h, R mu, R lam) {
  const M2<R> &F = h.F;
  const R I2 = F.a * F.a + F.b * F.b + F.c * F.c + F.d * F.d;
  return R(0.5) * mu * (I2 - R(2)) - mu * h.logJ + R(0.5) * lam * h.logJ * h.logJ;
}

// dF of a nodal direction (p0, p1, p2): dDs B
template <typename R>
__device__ __forceinline__ M2<R> dF_of(V2<R> p0, V2<R> p1, V2<R> p2, M2<R> B) {
  const M2<R> dDs = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
  return mat_mul(dDs, B);
}

// Directional second derivative d2Psi/dF2 [dF] as a matrix (for Hessian-vector products).
template <typename R>
__device__ __forceinline__ M2<R> nh_d2_dir(const NH<R> &h, M2<R> dF, R mu, R lam) {
  const M2<R> &G = h.G;
  const R GdF = G.a * dF.a + G.b * dF.b + G.c * dF.c + G.d * dF.d;
  const M2<R> GdFt = {G.a * dF.a + G.b * dF.b, G.a * dF.c + G.b * dF.d,
                      G.c * dF.a + G.d * dF.b, G.c * dF.c + G.d * dF.d};
  const M2<R> T = {GdFt.a * G.a + GdFt.b * G.c, GdFt.a * G.b + GdFt.b * G.d,
                   GdFt.c * G.a + GdFt.d * G.c, GdFt.c * G.b + GdFt.d * G.d};
  const R l = lam * GdF;
  return {mu * dF.a + l * G.a - h.c * T.a, mu * dF.b + l * G.b - h.c * T.b,
          mu * dF.c + l * G.c - h.c * T.c, mu * dF.d + l * G.d - h.c * T.d};
}

// Smallest positive s with 2 area(x + s p) = shrink 2 area(x), or BIG
// (reference implicit/kernels.py:304-341).
__device__ __forceinline__ double inversion_root(V2<double> e1, V2<double> e2, V2<double> q1,
                                                 V2<double> q2, double shrink) {
  const double a0 = cross(e1, e2);
  const double a1 = cross(e1, q2) + cross(q1, e2);
  const double a2 = cross(q1, q2);
  const double cc = a0 - shrink * a0;
  if (fabs(a2) < 1.0e-24) {
    if (fabs(a1) > 1.0e-24) {
      const double s = -cc / a1;
      if (s > 0.0) return s;
    }
    return BIG;
  }
  const double disc = a1 * a1 - 4.0 * a2 * cc;
  if (disc < 0.0) return BIG;
  const double sq = sqrt(disc);
  const double r1 = (-a1 - sq) / (2.0 * a2);
  const double r2 = (-a1 + sq) / (2.0 * a2);
  double s = BIG;
  if (r1 > 0.0) s = fmin(s, r1);
  if (r2 > 0.0) s = fmin(s, r2);
  return s;
}

// ------------------------------------------------------------------------------- barrier
// b(d) = -(d - dhat)^2 log(d / dhat) and its derivatives (reference core/math.py:113-129).
__device__ __forceinline__ double barrier_b(double d, double dhat) {
  const double u = d - dhat;
  return -(u * u) * log(d / dhat);
}
__device__ __forceinline__ double barrier_bp(double d, double dhat) {
  const double u = d - dhat;
  return -2.0 * u * log(d / dhat) - (u * u) / d;
}
__device__ __forceinline__ double barrier_bpp(double d, double dhat) {
  const double u = d - dhat;
  return -2.0 * log(d / dhat) - 4.0 * u / d + (u * u) / (d * d);
}

// ------------------------------------------------------------------------------- colliders
struct Collider {
  int kind;
  double cx, cy, rx, ry, rref, r, mu;
};

struct SdfHit {
  double d;
  V2<double> g;
  M2<double> H;
};

// Level set phi, grad phi and hess phi of a collider (reference core/sdf.py:45-137).
__device__ inline SdfHit sdf_eval(const Collider &c, V2<double> x) {
  SdfHit h;
  h.H = {0.0, 0.0, 0.0, 0.0};
  if (c.kind == 0) {
    const V2<double> diff = {x.x - c.cx, x.y - c.cy};
    const double dc = len(diff);
    const V2<double> n = {diff.x / dc, diff.y / dc};
    h.d = dc - c.rx;
    h.g = n;
    const double s = 1.0 / dc;
    h.H = {(1.0 - n.x * n.x) * s, (-n.x * n.y) * s, (-n.y * n.x) * s, (1.0 - n.y * n.y) * s};
  } else if (c.kind == 1) {
    const double ux = (x.x - c.cx) / c.rx, uy = (x.y - c.cy) / c.ry;
    const double u = ux * ux + uy * uy;
    const double su = sqrt(fmax(u, 1.0e-18));
    const V2<double> w = {(x.x - c.cx) / (c.rx * c.rx), (x.y - c.cy) / (c.ry * c.ry)};
    h.d = c.rref * (su - 1.0);
    const double k = c.rref / su;
    h.g = {k * w.x, k * w.y};
    const double iu = 1.0 / u;
    h.H = {k * (1.0 / (c.rx * c.rx) - iu * w.x * w.x), k * (-iu * w.x * w.y),
           k * (-iu * w.y * w.x), k * (1.0 / (c.ry * c.ry) - iu * w.y * w.y)};
  } else if (c.kind == 4) {
    h.d = c.rx * (x.x - c.cx) + c.ry * (x.y - c.cy);
    h.g = {c.rx, c.ry};
  } else if (c.kind == 2) {
    const double r = fmin(fmax(c.rref, 0.0), fmin(c.rx, c.ry));
    const double sx = x.x < c.cx ? -1.0 : 1.0;
    const double sy = x.y < c.cy ? -1.0 : 1.0;
    const double qx = sx * (x.x - c.cx) - (c.rx - r);
    const double qy = sy * (x.y - c.cy) - (c.ry - r);
    const double mx = fmax(qx, 0.0), my = fmax(qy, 0.0);
    const double ml = sqrt(mx * mx + my * my);
    h.d = ml + fmin(fmax(qx, qy), 0.0) - r;
    if (ml > 1.0e-12) {
      const double gx = mx / ml, gy = my / ml;
      h.g = {sx * gx, sy * gy};
      if (qx > 0.0 && qy > 0.0) {
        const double off = -gx * gy / ml;
        h.H = {gy * gy / ml, sx * sy * off, sx * sy * off, gx * gx / ml};
      }
    } else {
      h.g = qx > qy ? V2<double>{sx, 0.0} : V2<double>{0.0, sy};
    }
  } else {  // kind 3: rounded equilateral triangle, apex +y
    const double r3 = fmin(fmax(c.rref, 0.0), c.rx * 0.5);
    const double rin = c.rx - r3;
    const double a0 = 1.5707963267948966, da = 2.0943951023931953;
    V2<double> v[3];
    for (int k = 0; k < 3; ++k) v[k] = {c.cx + rin * cos(a0 + k * da), c.cy + rin * sin(a0 + k * da)};
    double dm = 0.0, sm = 0.0;
    V2<double> qm = {0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
      const V2<double> va = v[k], vb = v[(k + 1) % 3];
      const V2<double> ab = vb - va;
      const double t = dot(x - va, ab) / fmax(dot(ab, ab), 1.0e-30);
      const double tc = fmin(fmax(t, 0.0), 1.0);
      const V2<double> q = va + tc * ab;
      const double dd = len(x - q);
      const double isv = (t <= 0.0 || t >= 1.0) ? 1.0 : 0.0;
      if (k == 0 || dd < dm) {
        dm = dd;
        qm = q;
        sm = isv;
      }
    }
    bool in = true;
    for (int k = 0; k < 3; ++k) {
      const V2<double> e = v[(k + 1) % 3] - v[k];
      const V2<double> p = x - v[k];
      if (e.x * p.y - e.y * p.x < 0.0) in = false;
    }
    const double sgn = in ? -1.0 : 1.0;
    const double dl = fmax(dm, 1.0e-30);
    h.d = sgn * dm - r3;
    h.g = (sgn / dl) * (x - qm);
    if (sm > 0.5 && sgn > 0.0) {
      const V2<double> gg = h.g;
      const double s = 1.0 / dl;
      h.H = {(1.0 - gg.x * gg.x) * s, (-gg.x * gg.y) * s, (-gg.y * gg.x) * s, (1.0 - gg.y * gg.y) * s};
    }
  }
  return h;
}

// Barrier gradient and (clamped) diagonal of one boundary node against a collider
// (disc: implicit/kernels.py:132-157; level set: core/sdf.py:141-154). Returns false off the
// barrier's support.
__device__ inline bool collider_grad_diag(const Collider &c, V2<double> x, double dhat, double kappa,
                                          double h2, int clamp, V2<double> &g, V2<double> &dg) {
  if (c.kind == 0) {
    const V2<double> diff = {x.x - c.cx, x.y - c.cy};
    const double dc = len(diff);
    const double d = dc - c.r;
    if (d <= 0.0 || d >= dhat) return false;
    const V2<double> n = {diff.x / dc, diff.y / dc};
    const double bp = barrier_bp(d, dhat), bpp = barrier_bpp(d, dhat);
    g = (h2 * kappa * bp) * n;
    const double tang = bp / dc;
    double dx = kappa * (bpp * n.x * n.x + tang * (1.0 - n.x * n.x));
    double dy = kappa * (bpp * n.y * n.y + tang * (1.0 - n.y * n.y));
    if (clamp) { dx = fmax(dx, 0.0); dy = fmax(dy, 0.0); }
    dg = {h2 * dx, h2 * dy};
    return true;
  }
  const SdfHit s = sdf_eval(c, x);
  if (s.d <= 0.0 || s.d >= dhat) return false;
  const double bp = barrier_bp(s.d, dhat), bpp = barrier_bpp(s.d, dhat);
  g = (h2 * kappa * bp) * s.g;
  double dx = kappa * (bpp * s.g.x * s.g.x + bp * s.H.a);
  double dy = kappa * (bpp * s.g.y * s.g.y + bp * s.H.d);
  if (clamp) { dx = fmax(dx, 0.0); dy = fmax(dy, 0.0); }
  dg = {h2 * dx, h2 * dy};
  return true;
}

// Barrier p^T H p of one node; also returns the clearance (for clr_min) in `dist`
// (implicit/kernels.py:277-299, core/sdf.py:158-171).
__device__ inline double collider_pHp(const Collider &c, V2<double> x, V2<double> p, double dhat,
                                      double kappa, double h2, int clamp, double &dist) {
  double val;
  if (c.kind == 0) {
    const V2<double> diff = {x.x - c.cx, x.y - c.cy};
    const double dc = len(diff);
    const double d = dc - c.r;
    dist = d;
    if (d <= 0.0 || d >= dhat) return 0.0;
    const V2<double> n = {diff.x / dc, diff.y / dc};
    const double pn = dot(p, n);
    const V2<double> pt = p - pn * n;
    val = kappa * (barrier_bpp(d, dhat) * pn * pn + (barrier_bp(d, dhat) / dc) * dot(pt, pt));
  } else {
    const SdfHit s = sdf_eval(c, x);
    dist = s.d;
    if (s.d <= 0.0 || s.d >= dhat) return 0.0;
    const double gp = dot(s.g, p);
    const V2<double> Hp = {s.H.a * p.x + s.H.b * p.y, s.H.c * p.x + s.H.d * p.y};
    val = kappa * (barrier_bpp(s.d, dhat) * gp * gp + barrier_bp(s.d, dhat) * dot(p, Hp));
  }
  if (clamp) val = fmax(val, 0.0);
  return h2 * val;
}

__device__ inline double collider_energy(const Collider &c, V2<double> x, double dhat, double kappa,
                                         double h2) {
  double d;
  if (c.kind == 0) {
    d = len(V2<double>{x.x - c.cx, x.y - c.cy}) - c.r;
  } else {
    d = sdf_eval(c, x).d;
  }
  if (d <= 0.0 || d >= dhat) return 0.0;
  return h2 * kappa * barrier_b(d, dhat);
}

// Lagged friction frame at a node in contact: normal force, unit normal, tangential velocity
// (disc: implicit/kernels.py:661-690; level set: core/sdf.py:201-210).
struct FricFrame {
  double f_n;
  V2<double> n, vt;
  double vtm;
};

__device__ inline bool friction_frame(const Collider &c, V2<double> x, V2<double> xn, double dhat,
                                      double kappa, double inv_dt, FricFrame &f) {
  if (c.kind == 0) {
    const V2<double> diff = {x.x - c.cx, x.y - c.cy};
    const double dc = len(diff);
    const double d = dc - c.r;
    if (d <= 0.0 || d >= dhat) return false;
    f.n = {diff.x / dc, diff.y / dc};
    f.f_n = kappa * fabs(barrier_bp(d, dhat));
  } else {
    const SdfHit s = sdf_eval(c, x);
    if (s.d <= 0.0 || s.d >= dhat) return false;
    const double gn = len(s.g);
    const double ig = 1.0 / fmax(gn, 1.0e-30);
    f.n = {s.g.x * ig, s.g.y * ig};
    f.f_n = kappa * fabs(barrier_bp(s.d, dhat)) * gn;
  }
  const V2<double> v = inv_dt * (x - xn);
  const double vn = dot(v, f.n);
  f.vt = v - vn * f.n;
  f.vtm = len(f.vt);
  return true;
}

// Collider step cap: exact disc time of impact scaled by eta (implicit/kernels.py:345-371) or
// the conservative level-set bound eta phi / (-grad phi . p) (core/sdf.py:298-306). BIG if none.
__device__ inline double collider_ccd(const Collider &c, V2<double> x, V2<double> p, double eta) {
  if (c.kind == 0) {
    const V2<double> o = {x.x - c.cx, x.y - c.cy};
    const double a = dot(p, p);
    if (a < 1.0e-24) return BIG;
    const double bq = 2.0 * dot(o, p);
    const double cq = dot(o, o) - c.r * c.r;
    const double disc = bq * bq - 4.0 * a * cq;
    if (disc < 0.0) return BIG;
    const double sq = sqrt(disc);
    const double r1 = (-bq - sq) / (2.0 * a);
    const double r2 = (-bq + sq) / (2.0 * a);
    double s = BIG;
    if (r1 > 0.0) s = fmin(s, r1);
    if (r2 > 0.0) s = fmin(s, r2);
    return s < BIG ? eta * s : BIG;
  }
  const SdfHit s = sdf_eval(c, x);
  const double gp = dot(s.g, p);
  if (gp >= 0.0 || s.d <= 0.0) return BIG;
  return eta * s.d / (-gp);
}

// ------------------------------------------------------------------------------- IPC pairs
__device__ __forceinline__ double seg_dist(V2<double> p, V2<double> a, V2<double> b) {
  const V2<double> ab = b - a;
  double t = dot(p - a, ab) / fmax(dot(ab, ab), 1.0e-30);
  t = fmin(fmax(t, 0.0), 1.0);
  return len(p - a - t * ab);
}

// Additive CCD of a node against an edge (reference implicit/ipc_contact.py:103-127).
__device__ inline double accd(V2<double> p, V2<double> a, V2<double> b, V2<double> dp, V2<double> da,
                              V2<double> db, double g, double t_c) {
  const V2<double> m = (1.0 / 3.0) * (dp + da + db);
  dp = dp - m;
  da = da - m;
  db = db - m;
  const double lp = len(dp) + fmax(len(da), len(db));
  if (lp <= 1.0e-30) return t_c;
  double d = seg_dist(p, a, b);
  double t = 0.0;
  double tl = (d - g) / lp;
  for (int k = 0; k < 32; ++k) {
    const double tn = t + tl;
    if (tn >= t_c) return t_c;
    d = seg_dist(p + tn * dp, a + tn * da, b + tn * db);
    if (k > 0 && d < g) return t;
    t = tn;
    tl = 0.9 * d / lp;
  }
  return t;
}


// ------------------------------------------------------------------------------- rigid bodies
// Affine-body 6x6 block algebra (reference implicit/affine.py:61-147) and the 3x3 inverse of
// the rigid-mode coarse block (reference implicit/affine.py:465).
struct M66 {
  double m[6][6];
};

__device__ inline void affine_block(M66 &H, double A11, double A12, double A21, double A22, double Mu,
                                    double La, double h2) {
  const int idx[4] = {1, 2, 4, 5};
  const NH<double> h = nh_eval(M2<double>{A11, A12, A21, A22}, Mu, La);
  for (int k = 0; k < 4; ++k) {
    M2<double> E = {0.0, 0.0, 0.0, 0.0};
    if (k == 0) E.a = 1.0;
    else if (k == 1) E.b = 1.0;
    else if (k == 2) E.c = 1.0;
    else E.d = 1.0;
    const M2<double> col = nh_d2_dir(h, E, Mu, La);
    const double cv[4] = {col.a, col.b, col.c, col.d};
    for (int r = 0; r < 4; ++r) H.m[idx[r]][idx[k]] += h2 * cv[r];
  }
  const double a = A11 + A22, b = A21 - A12;
  const double n = sqrt(a * a + b * b);
  if (n > 1.0e-12) {
    const double s = 0.7071067811865476 / n;
    const double t[4] = {-b * s, -a * s, a * s, -b * s};
    const M2<double> Tm = {t[0], t[1], t[2], t[3]};
    const double lam_t = nh_quad(h, Tm, Mu, La);
    if (lam_t < 0.0)
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) H.m[idx[r]][idx[c]] -= h2 * lam_t * t[r] * t[c];
  }
}

__device__ inline void cholesky6(const M66 &H, M66 &L) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) L.m[i][j] = 0.0;
  bool ok = true;
  for (int j = 0; j < 6; ++j) {
    double s = H.m[j][j];
    for (int k = 0; k < j; ++k) s -= L.m[j][k] * L.m[j][k];
    if (s <= 1.0e-9 * fabs(H.m[j][j]) || s <= 0.0) ok = false;
    if (ok) {
      const double d = sqrt(s);
      L.m[j][j] = d;
      for (int i = j + 1; i < 6; ++i) {
        double v = H.m[i][j];
        for (int k = 0; k < j; ++k) v -= L.m[i][k] * L.m[j][k];
        L.m[i][j] = v / d;
      }
    }
  }
  if (!ok) {
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) L.m[i][j] = 0.0;
    for (int j = 0; j < 6; ++j) L.m[j][j] = sqrt(fmax(H.m[j][j], 1.0e-30));
  }
}

__device__ inline void cholesky6_solve(const M66 &L, const double *b, double *x) {
  double y[6];
  for (int i = 0; i < 6; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L.m[i][k] * y[k];
    y[i] = s / L.m[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= L.m[k][i] * x[k];
    x[i] = s / L.m[i][i];
  }
}

__device__ inline void invert33(const double *H, double *Hi) {
  const double a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7], i = H[8];
  const double A = e * i - f * h, Bc = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * Bc + c * C;
  const double id = 1.0 / det;
  Hi[0] = A * id;
  Hi[1] = -(b * i - c * h) * id;
  Hi[2] = (b * f - c * e) * id;
  Hi[3] = Bc * id;
  Hi[4] = (a * i - c * g) * id;
  Hi[5] = -(a * f - c * d) * id;
  Hi[6] = C * id;
  Hi[7] = -(a * h - b * g) * id;
  Hi[8] = (a * e - b * d) * id;
}
// ------------------------------------------------------------------ springs and pins
// E = k/2 (|d| - L0)^2, d = x_a - x_b (reference implicit/actuation.py:368-392): the tension
// k (|d| - L0), the tangential stiffness k (1 - L0/|d|) clamped at 0, the length and the unit
// direction of d.
struct SpringGeo {
  V2<double> n;
  double tension, tang, len;
};
__device__ __forceinline__ SpringGeo spring_geo(V2<double> xa, V2<double> xb, double k, double L0) {
  const V2<double> d = xa - xb;
  SpringGeo g;
  g.len = sqrt(dot(d, d) + 1.0e-30);
  g.tension = k * (g.len - L0);
  g.tang = fmax(k * (1.0 - L0 / g.len), 0.0);
  g.n = {d.x / g.len, d.y / g.len};
  return g;
}
// Diagonal of the spring's Hessian at either end, h2 (k n n^T + tang (I - n n^T)).
__device__ __forceinline__ V2<double> spring_diag(const SpringGeo &g, double k, double h2) {
  return {h2 * (k * g.n.x * g.n.x + g.tang * (1.0 - g.n.x * g.n.x)),
          h2 * (k * g.n.y * g.n.y + g.tang * (1.0 - g.n.y * g.n.y))};
}
// p^T H p of the spring along the relative direction dp = p_a - p_b (p_a for a world spring).
__device__ __forceinline__ double spring_pHp(const SpringGeo &g, double k, double h2, V2<double> dp) {
  const double along = dot(dp, g.n);
  return h2 * (k * along * along + g.tang * (dot(dp, dp) - along * along));
}

}  // namespace d2f
