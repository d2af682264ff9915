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
                          int32_t B, const uint8_t *excl, const int64_t *excl_off, double radius,
                          double stiffness, double h2, int32_t clamp, double *g, double *dg) {
  NodeArgs a{};
  a.x = x;
  a.pdir = v;
  a.bnd = bnd;
  a.bnd_start = bnd_start;
  a.B = B;
  a.max_nb = 0;  // sizes unknown on the host: partners read from global memory
  a.excl = excl;
  a.excl_off = excl_off;
  a.k = stiffness;
  a.r = radius;
  a.h2 = h2;
  a.clamp = clamp;
  a.g = g;
  a.dg = dg;
  return a;
}

extern "C" int d2f_node_grad_diag_f64(const double *x, const int32_t *bnd, const int32_t *bnd_start, int32_t B,
                                      const uint8_t *excl, const int64_t *excl_off, double radius,
                                      double stiffness, double h2, int32_t clamp, double *g, double *dg,
                                      void *stream) {
  D2F_CHECK(B >= 0 && radius >= 0.0, "d2f_node_grad_diag_f64: bad argument");
  D2F_CHECK(B == 0 || (x && bnd && bnd_start && excl && excl_off && g && dg),
            "d2f_node_grad_diag_f64: null pointer");
  return launch_node_env(0, node_args(x, nullptr, bnd, bnd_start, B, excl, excl_off, radius, stiffness, h2,
                                      clamp, g, dg),
                         (hipStream_t)stream);
}

extern "C" int d2f_node_hvp_f64(const double *x, const double *v, const int32_t *bnd, const int32_t *bnd_start,
                                int32_t B, const uint8_t *excl, const int64_t *excl_off, double radius,
                                double stiffness, double h2, double *y, void *stream) {
  D2F_CHECK(B >= 0 && radius >= 0.0, "d2f_node_hvp_f64: bad argument");
  D2F_CHECK(B == 0 || (x && v && bnd && bnd_start && excl && excl_off && y), "d2f_node_hvp_f64: null pointer");
  return launch_node_env(3, node_args(x, v, bnd, bnd_start, B, excl, excl_off, radius, stiffness, h2, 0, y,
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
