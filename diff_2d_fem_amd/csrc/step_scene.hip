// step_scene.hip — the scene-persistent path: one workgroup owns one environment and runs the
// whole implicit step (every PNCG iteration of reference implicit/solver.py:252-451) inside one
// launch. Per-environment scalars live in LDS; reductions are fixed-order block reductions. The
// elastic scatter adds into per-node LDS accumulators with float64 LDS atomics when the scene's
// nodes fit (the order of those additions, and with it the last bits of a sum, can differ
// between runs), else colour by colour with plain read-modify-write of the node arrays (no two
// triangles of a colour share a node). Contact terms accumulate per boundary node in LDS; the
// node-edge broadphase is a per-scene uniform grid and Verlet pair list built in LDS. An
// environment leaves the loop at the first convergence check it is converged at — the state it
// leaves is the one the reference's global loop would leave it in, because a converged
// environment's positions no longer change (implicit/kernels.py:375-434).
//
// Template R is the element arithmetic: double (float64 path) or float (mixed: float32 element
// stress and curvature from a displacement gradient formed in float64; float64 state, dot
// products and energies).
#include <algorithm>
#include <atomic>
#include <mutex>

#include "common.h"
#include "node_contact.h"
#include "scene_dev.h"

namespace d2f {

using SH = d2f_scene;  // host view (kernel argument, launch)
// device views, read in place from the kernel-argument segment (constant address space):
// uniform fields become scalar loads and array pointers are known to be global
using S = const __attribute__((address_space(4))) SceneG;
using KPm = const __attribute__((address_space(4))) d2f_params;
using LdsD = __attribute__((address_space(3))) double;
using LdsI = __attribute__((address_space(3))) int;
using LdsU = __attribute__((address_space(3))) uint32_t;
using GD = D2F_GLOBAL double;
using GCD = const D2F_GLOBAL double;
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
// atomics on LDS: the explicit cast to a generic pointer is folded back to an LDS operation
__device__ __forceinline__ void lds_add(LdsD *a, double v) { atomicAdd((double *)a, v); }
__device__ __forceinline__ int lds_add(LdsI *a, int v) { return atomicAdd((int *)a, v); }
using Pm = d2f_params;
constexpr int MAXC = D2F_MAX_COLORS;
constexpr int RED_K = 20;  // widest block reduction

// ----------------------------------------------------------------------------- LDS state
struct SceneScalars {
  double gPg, g_p, g_Py, y_p, y_Py, beta, gTp, pHp, alpha, p_inf, clr_min, dE, dE0;
  double clr_ipc, q_inf, sweep_beta, Eprev, Ener, tc;
  double gx0, gy0, gh;          // broadphase grid origin and cell size
  int gnx, gny, active, iters, pending, stop;
  int npairs, plist, excl_lds;  // candidate pairs of the last build, list valid, bitset in LDS
  int stale;                    // stamp of the last contact_prepare that found a node past skin/2
};
using LdsScalars = __attribute__((address_space(3))) SceneScalars;

// A collider read from the kernel arguments.
__device__ __forceinline__ Collider to_collider_k(const __attribute__((address_space(4))) d2f_collider &c) {
  return Collider{c.kind, c.cx, c.cy, c.rx, c.ry, c.rref, c.r, c.mu};
}

// Cross-lane steps of one double in DPP form (two 32-bit halves moved by the VALU, no LDS):
// CTRL 0xB1 swaps neighbouring lanes, 0x4E lanes two apart, 0x124 / 0x128 rotate each row of
// 16 lanes by 4 / 8.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), l), hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
// Reduction over each row of 16 lanes, the row's result in all of its lanes (all lanes active).
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x124>(v);
  return v + dpp_f64<0x128>(v);
}
__device__ __forceinline__ double row_min(double v) {
  v = fmin(v, dpp_f64<0xB1>(v));
  v = fmin(v, dpp_f64<0x4E>(v));
  v = fmin(v, dpp_f64<0x124>(v));
  return fmin(v, dpp_f64<0x128>(v));
}
// Whole-wave results, uniform over the wave: the four row results read back as scalars.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = row_sum(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ double wave_min_dpp(double v) {
  v = row_min(v);
  return fmin(fmin(readlane_f64(v, 0), readlane_f64(v, 16)), fmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Workgroup reductions, called by every thread of the workgroup with every lane active; each
// thread gets the result.
// * K <= 2: each wave reduces its values completely; lane 0 of each wave parks them in red[];
//   then lane w of every wave picks up wave w's value and one row reduction combines them.
// * K >= 3: each wave only reduces its rows of 16 lanes, and lane 0 of each row parks the row
//   result (P = 4 NW partials per value). Wave w then combines the partials of values
//   w, w + NW, ... and parks the totals, which every thread reads back. A wave issues about
//   K row reductions plus K / NW combinations instead of K full reductions plus K
//   combinations, for one more barrier.
template <int TPB>
struct Block {
  static constexpr int NW = TPB / WAVE;
  static constexpr int P = 4 * NW;                   // row partials per value
  static constexpr int RED_SIZE = RED_K * P + RED_K;  // doubles of LDS: partials, then totals
  static_assert(NW <= 16, "one wave combines the row partials");
  LdsD *red;  // [RED_SIZE]

  template <bool MIN>
  static __device__ __forceinline__ double row_op(double v) { return MIN ? row_min(v) : row_sum(v); }
  template <bool MIN>
  static __device__ __forceinline__ double op(double a, double b) { return MIN ? fmin(a, b) : a + b; }

  // v[k] summed for k < NS, minimum for k >= NS
  template <int NS, int K>
  __device__ __forceinline__ void reduce(double (&v)[K]) {
    static_assert(K <= RED_K, "reduction wider than the LDS scratch");
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    if constexpr (K <= 2) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = k >= NS ? wave_min_dpp(v[k]) : wave_sum_dpp(v[k]);
      __syncthreads();
      if (l == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[w * K + k] = v[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; ++k)
        v[k] = k >= NS ? readlane_f64(row_min(l < NW ? double(red[l * K + k]) : __builtin_huge_val()), 0)
                       : readlane_f64(row_sum(l < NW ? double(red[l * K + k]) : 0.0), 0);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = k >= NS ? row_min(v[k]) : row_sum(v[k]);
      __syncthreads();
      if ((l & 15) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * P + 4 * w + (l >> 4)] = v[k];
      __syncthreads();
      LdsD *tot = red + RED_K * P;
      for (int k = w; k < K; k += NW) {   // uniform over the wave
        if (k >= NS) tot_row<true>(tot, k, l);
        else tot_row<false>(tot, k, l);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = tot[k];
    }
  }
  // the total of value k from its P row partials, parked by lane 0
  template <bool MIN>
  __device__ __forceinline__ void tot_row(LdsD *tot, int k, int l) {
    const double idle = MIN ? __builtin_huge_val() : 0.0;
    const double t = row_op<MIN>(l < P ? double(red[k * P + l]) : idle);
    double r = readlane_f64(t, 0);
    if constexpr (P > 16) r = op<MIN>(r, readlane_f64(t, 16));
    if constexpr (P > 32) r = op<MIN>(r, op<MIN>(readlane_f64(t, 32), readlane_f64(t, 48)));
    if (l == 0) tot[k] = r;
  }
  template <int K>
  __device__ void sum(double (&v)[K]) { reduce<K>(v); }
  // component-wise minimum
  template <int K>
  __device__ void min(double (&v)[K]) { reduce<0>(v); }
  __device__ double sum1(double a) {
    double v[1] = {a};
    sum<1>(v);
    return v[0];
  }
  __device__ double min1(double a) {
    double v[1] = {a};
    min<1>(v);
    return v[0];
  }
  __device__ double max1(double a) { return -min1(-a); }
  // exclusive prefix sum of one int per thread
  __device__ int scan_excl(int v, int &total) {
    int incl = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, WAVE);
      if ((int)(threadIdx.x % WAVE) >= o) incl += t;
    }
    LdsI *ired = reinterpret_cast<LdsI *>(red);
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    __syncthreads();
    if (l == WAVE - 1) ired[w] = incl;
    __syncthreads();
    int off = 0;
    total = 0;
    for (int i = 0; i < NW; ++i) {
      if (i < w) off += ired[i];
      total += ired[i];
    }
    return off + incl - v;
  }
};

// ----------------------------------------------------------------------------- element access
template <typename R>
struct Elem {
  int i0, i1, i2;
  M2<R> B;
  R Amu, Alam;
};

// The element arrays of one precision, read once from the scene struct into registers: the
// struct lives in LDS, and a loop with LDS atomics in it would otherwise re-read every pointer
// from there on each trip.
template <typename R>
struct ElemPtrs {
  const D2F_GLOBAL int32_t *tri;
  const D2F_GLOBAL R *Bm, *Amu, *Alam;
};
template <typename R>
__device__ __forceinline__ ElemPtrs<R> elem_ptrs(S &s);
template <>
__device__ __forceinline__ ElemPtrs<double> elem_ptrs<double>(S &s) {
  return {s.sc_tri, s.sc_Bm, s.sc_Amu, s.sc_Alam};
}
template <>
__device__ __forceinline__ ElemPtrs<float> elem_ptrs<float>(S &s) {
  return {s.sc_tri, s.sc_Bm32, s.sc_Amu32, s.sc_Alam32};
}
template <typename R>
__device__ __forceinline__ Elem<R> load_elem(const ElemPtrs<R> &q, int t) {
  Elem<R> el;
  el.i0 = q.tri[3 * t];
  el.i1 = q.tri[3 * t + 1];
  el.i2 = q.tri[3 * t + 2];
  el.B = {q.Bm[4 * t], q.Bm[4 * t + 1], q.Bm[4 * t + 2], q.Bm[4 * t + 3]};
  el.Amu = q.Amu[t];
  el.Alam = q.Alam[t];
  return el;
}

template <typename R>
__device__ __forceinline__ Elem<R> load_elem(S &s, int t);

template <>
__device__ __forceinline__ Elem<double> load_elem<double>(S &s, int t) {
  Elem<double> el;
  el.i0 = s.sc_tri[3 * t];
  el.i1 = s.sc_tri[3 * t + 1];
  el.i2 = s.sc_tri[3 * t + 2];
  const dvec4 b = reinterpret_cast<const D2F_GLOBAL dvec4 *>(s.sc_Bm)[t];
  el.B = {b.x, b.y, b.z, b.w};
  el.Amu = s.sc_Amu[t];
  el.Alam = s.sc_Alam[t];
  return el;
}
template <>
__device__ __forceinline__ Elem<float> load_elem<float>(S &s, int t) {
  Elem<float> el;
  el.i0 = s.sc_tri[3 * t];
  el.i1 = s.sc_tri[3 * t + 1];
  el.i2 = s.sc_tri[3 * t + 2];
  const fvec4 b = reinterpret_cast<const D2F_GLOBAL fvec4 *>(s.sc_Bm32)[t];
  el.B = {b.x, b.y, b.z, b.w};
  el.Amu = s.sc_Amu32[t];
  el.Alam = s.sc_Alam32[t];
  return el;
}

// Neo-Hookean state of an element from float64 positions. float64: F = Ds Bm.
// Mixed, for the gradient (ROT = true): F is formed in float64 and split into a rotation and
// a symmetric stretch, F = Q S (2-D polar decomposition: Q turns by atan2(F21 - F12,
// F11 + F22), which makes Q^T F symmetric). Only the stretch's displacement gradient S - I,
// small with the strain whatever the rotation, is rounded to float; the float stress is formed
// from it without cancellation (nh_eval_strain) and turned back by Q (the material is
// isotropic: P(Q S) = Q P(S)). The float stress so keeps its relative accuracy for small
// strains under any rotation. Q's normalisation is a float reciprocal square root refined by
// one float64 Newton step (relative error ~1e-14; Q is applied both ways, so any error in it
// only rescales F by that much).
// Mixed, for the curvature passes (ROT = false): edge vectors formed in float64, then F in
// float; their accuracy sets step lengths only.
template <typename R, bool ROT = true>
__device__ __forceinline__ NH<R> elem_F(const Elem<R> &el, V2<double> p0, V2<double> p1, V2<double> p2) {
  if constexpr (sizeof(R) == 4) {
    if constexpr (!ROT) {
      const M2<float> Ds = {float(p1.x - p0.x), float(p2.x - p0.x), float(p1.y - p0.y), float(p2.y - p0.y)};
      return nh_eval(mat_mul(Ds, el.B), el.Amu, el.Alam);
    } else {
      const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
      const M2<double> Bd = {double(el.B.a), double(el.B.b), double(el.B.c), double(el.B.d)};
      const M2<double> Fd = mat_mul(Ds, Bd);
      const double tc = Fd.a + Fd.d, ts = Fd.c - Fd.b;
      const double r2 = tc * tc + ts * ts;
      double cq = 1.0, sq = 0.0;
      if (r2 > 1.0e-30) {
        const double y0 = double(rsqrtf(float(r2)));
        const double ir = y0 * (1.5 - 0.5 * r2 * y0 * y0);
        cq = tc * ir;
        sq = ts * ir;
      }
      // S = Q^T F
      const M2<float> H = {float(cq * Fd.a + sq * Fd.c - 1.0), float(cq * Fd.b + sq * Fd.d),
                           float(cq * Fd.c - sq * Fd.a), float(cq * Fd.d - sq * Fd.b - 1.0)};
      NH<float> h = nh_eval_strain(H, el.Amu, el.Alam);
      const float c = float(cq), s = float(sq);
      const auto turn = [c, s](M2<float> A) -> M2<float> {
        return {c * A.a - s * A.c, c * A.b - s * A.d, s * A.a + c * A.c, s * A.b + c * A.d};
      };
      h.F = turn(h.F);
      h.G = turn(h.G);
      h.D = turn(h.D);
      return h;
    }
  } else {
    const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
    return nh_eval(mat_mul(Ds, el.B), el.Amu, el.Alam);
  }
}

// Elastic energy density times area of an element, in float64 whatever the element arithmetic
// (the line search compares energies; the float constants are promoted).
template <typename R>
__device__ __forceinline__ double elem_psi(const Elem<R> &el, V2<double> p0, V2<double> p1, V2<double> p2) {
  const M2<double> Ds = {p1.x - p0.x, p2.x - p0.x, p1.y - p0.y, p2.y - p0.y};
  const M2<double> Bd = {double(el.B.a), double(el.B.b), double(el.B.c), double(el.B.d)};
  const double mu = double(el.Amu), lam = double(el.Alam);
  return nh_psi(nh_eval(mat_mul(Ds, Bd), mu, lam), mu, lam);
}

template <typename R>
__device__ __forceinline__ V2<double> up(V2<R> v) {
  return {double(v.x), double(v.y)};
}

// ----------------------------------------------------------------------------- the kernel
// FULL: the launch uses a term beyond elasticity, colliders, IPC contact and affine bodies
// (tendons, springs, pressure chambers, rigid modes, "node" contact or the line search).
// Without FULL their ranges are the constant 0 and their flags false, so their passes compile
// away: the kernel's code then weighs less on the register allocation and on the instruction
// cache.
template <typename R, int TPB, bool FULL>
struct Step {
  S &s;
  KPm &p;
  LdsScalars &sc;
  __device__ __forceinline__ bool node_on() const { return FULL && p.node_on; }
  __device__ __forceinline__ bool line_search() const { return FULL && p.line_search; }
  Block<TPB> blk;
  LdsD *bpos;     // 2 nb: boundary positions
  LdsD *bacc;     // 4 nb: boundary accumulators (grad x, grad y, diag x, diag y)
  LdsD *bref;     // 2 nb: boundary positions the grid and pair list were built from
  GD *gcur;       // this iteration's gradient (s.grad or s.grad_prev)
  GD *gprev;      // the previous iteration's
  LdsD *nacc;     // 4 x max_nn, component-major: grad x, grad y, diag x, diag y of each node
  // The scene's affine groups in LDS for the whole step, GST doubles each: coordinates q (6),
  // their line-search start q0 (6), direction pq (6), preconditioned gradient zq (6), Mu, La.
  // Loaded from ag_q, ag_pq, ag_zq, ag_Mu, ag_La at the start, written back at the end.
  static constexpr int GST = 26;
  LdsD *gst;
  __device__ __forceinline__ LdsD *gq_(int g) const { return gst + GST * (g - g0); }
  __device__ __forceinline__ LdsD *gq0_(int g) const { return gq_(g) + 6; }
  __device__ __forceinline__ LdsD *gpq_(int g) const { return gq_(g) + 12; }
  __device__ __forceinline__ LdsD *gzq_(int g) const { return gq_(g) + 18; }
  __device__ __forceinline__ double gMu_(int g) const { return gq_(g)[24]; }
  __device__ __forceinline__ double gLa_(int g) const { return gq_(g)[25]; }
  LdsD *xc;       // 2 nn: the scene's positions (when xc_lds), written through with s.x
  bool xcon;
  LdsI *cstart;   // max_cells + 1
  LdsI *citem;    // nb
  LdsI *ccur;     // max_cells
  LdsI *lbnd;     // the scene's boundary nodes (bnd[b0 ..]), copied once per step
  int prep_stamp = 0;   // calls of contact_prepare so far (the same in every thread)
  LdsU *xbits;    // exclusion bitset (excl_bits_lds words)
  LdsI *plist;    // max_pairs packed (edge - k0) << 16 | jp
  int e, tid;
  int n0, n1, b0, nb, k0, k1, t0, t1, c0, bc0, bc1, g0, g1, r0, r1, pe0, pe1, td0, td1, sp0, sp1;
  int64_t eoff;

  __device__ __forceinline__ Step(S &s_, KPm &p_, LdsScalars &sc_, LdsD *lds)
      : s(s_), p(p_), sc(sc_) {
    e = s.sc_list ? s.sc_list[blockIdx.x] : blockIdx.x;
    tid = threadIdx.x;
    gcur = s.grad;
    gprev = s.grad_prev;
    blk.red = lds;
    bpos = lds + Block<TPB>::RED_SIZE;
    bacc = bpos + 2 * s.max_nb;
    bref = bacc + 4 * s.max_nb;
    gst = bref + 2 * s.max_nb;
    nacc = gst + GST * s.max_ag;
    xc = nacc + (s.nacc_lds ? 4 * s.max_nn : 0);
    cstart = reinterpret_cast<LdsI *>(xc + (s.xc_lds ? 2 * s.max_nn : 0));
    citem = cstart + s.max_cells + 1;
    ccur = citem + s.max_nb;
    lbnd = ccur + s.max_cells;
    xbits = reinterpret_cast<LdsU *>(lbnd + s.max_nb);
    plist = reinterpret_cast<LdsI *>(xbits + s.excl_bits_lds);
    n0 = s.node_start[e];
    n1 = s.node_start[e + 1];
    b0 = s.bnd_start[e];
    nb = s.bnd_start[e + 1] - b0;
    k0 = s.n_e > 0 ? s.e_start[e] : 0;
    k1 = s.n_e > 0 ? s.e_start[e + 1] : 0;
    t0 = s.sc_tri_start[e];
    t1 = s.sc_tri_start[e + 1];
    c0 = e * (MAXC + 1);
    bc0 = s.n_bc > 0 ? s.bc_start[e] : 0;
    bc1 = s.n_bc > 0 ? s.bc_start[e + 1] : 0;
    g0 = s.n_ag > 0 ? s.ag_start[e] : 0;
    g1 = s.n_ag > 0 ? s.ag_start[e + 1] : 0;
    r0 = FULL && s.n_rg > 0 ? s.rg_start[e] : 0;
    r1 = FULL && s.n_rg > 0 ? s.rg_start[e + 1] : 0;
    pe0 = FULL && s.n_pe > 0 ? s.pe_start[e] : 0;
    pe1 = FULL && s.n_pe > 0 ? s.pe_start[e + 1] : 0;
    td0 = FULL && s.n_td > 0 ? s.td_start[e] : 0;
    td1 = FULL && s.n_td > 0 ? s.td_start[e + 1] : 0;
    sp0 = FULL && s.n_sp > 0 ? s.sp_start[e] : 0;
    sp1 = FULL && s.n_sp > 0 ? s.sp_start[e + 1] : 0;
    eoff = s.excl_bits_off ? s.excl_bits_off[e] : 0;
    xcon = s.xc_lds && n1 - n0 <= s.max_nn;
  }

  // Position of node i (of this scene): from the LDS copy when it is kept, else from s.x.
  __device__ __forceinline__ V2<double> X(int i) const {
    if (xcon) {
      const int l = 2 * (i - n0);
      return {xc[l], xc[l + 1]};
    }
    return ld2(s.x, i);
  }
  // Write node i's position to s.x and, when it is kept, to the LDS copy.
  __device__ __forceinline__ void xput(int i, V2<double> v) {
    st2(s.x, i, v);
    if (xcon) {
      const int l = 2 * (i - n0);
      xc[l] = v.x;
      xc[l + 1] = v.y;
    }
  }

  __device__ __forceinline__ double fr(int i) const { return s.free_[i] ? 1.0 : 0.0; }

  // A loop over the scene's nodes in batches of NB per thread: every input of a batch is loaded
  // before any is used, so the batch's memory latencies overlap. Each node's inputs and
  // arithmetic, and the order in which a thread visits its nodes, are those of a loop with one
  // node per trip.
  // `t` is the thread's slot among the `nt` threads that share the range.
  template <int NB, typename Load, typename Use>
  __device__ __forceinline__ void range_batches(int lo, int hi, int t, int nt, Load &&load, Use &&use) {
    using In = decltype(load(0));
    for (int b = lo + t; b < hi; b += NB * nt) {
      In in[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = b + u * nt;
        if (i < hi) in[u] = load(i);
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = b + u * nt;
        if (i < hi) use(i, in[u]);
      }
    }
  }
  template <int NB, typename Load, typename Use>
  __device__ __forceinline__ void range_batches(int lo, int hi, Load &&load, Use &&use) {
    range_batches<NB>(lo, hi, tid, TPB, load, use);
  }
  template <int NB, typename Load, typename Use>
  __device__ __forceinline__ void node_batches(Load &&load, Use &&use) {
    range_batches<NB>(n0, n1, load, use);
  }

  // ------------------------------------------------------------------ IPC broadphase grid
  // The boundary positions in LDS (current) are compared with those the pair list was built
  // from; when some node has moved more than half the skin, or there is no valid list, the
  // grid and the list are rebuilt. A pair within reach now was within reach + skin at the
  // build (each end moved at most skin / 2), so the list holds every pair a pass needs.
  // Contact distance the broadphase serves: the IPC reach or the "node" contact radius.
  __device__ __forceinline__ double creach() const { return node_on() ? p.node_r : p.ipc_reach; }

  // A node that moved more than half the skin marks the list stale with this call's stamp
  // (no reduction needed: one barrier, then every thread reads the same flag).
  __device__ __forceinline__ void contact_prepare() {
    const bool have = sc.plist;
    const int stamp = ++prep_stamp;
    for (int j = tid; j < nb; j += TPB) {
      const V2<double> x = X(lbnd[j]);
      bpos[2 * j] = x.x;
      bpos[2 * j + 1] = x.y;
      if (have && 2.0 * len(V2<double>{x.x - bref[2 * j], x.y - bref[2 * j + 1]}) > p.ipc_skin) sc.stale = stamp;
    }
    __syncthreads();
    if (!have || sc.stale == stamp) {
      build_grid();
      if (node_on()) {  // node pairs are found in the grid itself; plist marks it current
        if (tid == 0) sc.plist = 1;
        __syncthreads();
      } else {
        build_pairs();
      }
    }
  }

  // A uniform grid over the boundary positions in LDS (cell >= reach + skin, so a pair within
  // that distance is found among the cells the edge's box +- reach + skin overlaps).
  __device__ __forceinline__ void build_grid() {
    double bb[4] = {BIG, BIG, BIG, BIG};  // min x, min y, -max x, -max y
    for (int j = tid; j < nb; j += TPB) {
      const V2<double> x = {bpos[2 * j], bpos[2 * j + 1]};
      bref[2 * j] = x.x;
      bref[2 * j + 1] = x.y;
      bb[0] = fmin(bb[0], x.x);
      bb[1] = fmin(bb[1], x.y);
      bb[2] = fmin(bb[2], -x.x);
      bb[3] = fmin(bb[3], -x.y);
    }
    blk.template min<4>(bb);
    if (tid == 0) {
      double h = creach() + p.ipc_skin;
      const double wx = -bb[2] - bb[0], wy = -bb[3] - bb[1];
      int nx = 1, ny = 1;
      if (wx >= 0.0 && wy >= 0.0 && wx < 1.0e12 && wy < 1.0e12) {  // finite positions
        for (int k = 0; k < 400; ++k) {
          nx = (int)fmin(wx / h + 1.0, 1.0e9);
          ny = (int)fmin(wy / h + 1.0, 1.0e9);
          if ((long)nx * ny <= s.max_cells) break;
          h *= 1.25;
        }
        if ((long)nx * ny > s.max_cells) nx = ny = 1;
      }
      sc.gx0 = bb[0];
      sc.gy0 = bb[1];
      sc.gh = h;
      sc.gnx = nx;
      sc.gny = ny;
    }
    __syncthreads();
    const int nc = sc.gnx * sc.gny;
    for (int c = tid; c <= nc; c += TPB) cstart[c] = 0;
    __syncthreads();
    for (int j = tid; j < nb; j += TPB) lds_add(&cstart[cell_of(bpos[2 * j], bpos[2 * j + 1])], 1);
    __syncthreads();
    // exclusive scan of the counts, in chunks per thread
    const int chunk = (nc + TPB - 1) / TPB;
    const int lo = tid * chunk, hi = min(lo + chunk, nc);
    int loc = 0;
    for (int c = lo; c < hi; ++c) loc += cstart[c];
    int total;
    int off = blk.scan_excl(loc, total);
    __syncthreads();
    for (int c = lo; c < hi; ++c) {
      const int v = cstart[c];
      cstart[c] = off;
      off += v;
    }
    if (tid == 0) cstart[nc] = total;
    __syncthreads();
    // place items by an atomic cursor, then sort each cell so that the candidate order (and
    // with it every per-edge partial sum) does not depend on the placement order
    for (int c = tid; c < nc; c += TPB) ccur[c] = cstart[c];
    __syncthreads();
    for (int j = tid; j < nb; j += TPB) {
      const int c = cell_of(bpos[2 * j], bpos[2 * j + 1]);
      citem[lds_add(&ccur[c], 1)] = j;
    }
    __syncthreads();
    for (int c = tid; c < nc; c += TPB)
      for (int q = cstart[c] + 1; q < cstart[c + 1]; ++q) {
        const int v = citem[q];
        int r = q - 1;
        while (r >= cstart[c] && citem[r] > v) {
          citem[r + 1] = citem[r];
          --r;
        }
        citem[r + 1] = v;
      }
    __syncthreads();
  }

  __device__ __forceinline__ int cell_of(double x, double y) const {
    int ix = (int)((x - sc.gx0) / sc.gh);
    int iy = (int)((y - sc.gy0) / sc.gh);
    ix = ix < 0 ? 0 : (ix >= sc.gnx ? sc.gnx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= sc.gny ? sc.gny - 1 : iy);
    return iy * sc.gnx + ix;
  }

  // ------------------------------------------------------------------ IPC pair terms
  // One barrier-and-friction term (reference implicit/ipc_contact.py:131-171); jp, ja, jb are
  // scene-local boundary indices. mode 0 accumulates into bacc (LDS atomics), 1 returns the
  // p^T H p share, 2 the energy share.
  __device__ __forceinline__ double term(int mode, double coef, int jp, int ja, int jb, int ip, int ia, int ib,
                         double wa, double wb, double d, V2<double> n, double ddd, double mu,
                         GCD *pdir, double h2) {
    const double dhat = p.dhat, kappa = p.ipc_kappa;
    const double bp = barrier_bp(d, dhat);
    const double f_n = -kappa * bp;
    const V2<double> tn = {-n.y, n.x};
    const V2<double> u = (X(ip) - ld2(s.xn, ip)) - wa * (X(ia) - ld2(s.xn, ia)) -
                         wb * (X(ib) - ld2(s.xn, ib));
    const double ut = dot(u, tn);
    const double vt = fabs(ut) * p.inv_dt;
    const double kt = mu * f_n * p.inv_dt / fmax(vt, p.epsv);
    if (mode == 0) {
      const double bpp = barrier_bpp(d, dhat);
      const V2<double> g = h2 * ((-f_n) * n + (mu * f_n * (ut * p.inv_dt / fmax(vt, p.epsv))) * tn);
      const V2<double> hd = {h2 * (kappa * bpp * n.x * n.x + kt * tn.x * tn.x),
                             h2 * (kappa * bpp * n.y * n.y + kt * tn.y * tn.y)};
      const double fp = fr(ip), fa = fr(ia), fb = fr(ib);
      LdsD *A = bacc + 4 * jp;
      lds_add(A, coef * fp * g.x);
      lds_add(A + 1, coef * fp * g.y);
      lds_add(A + 2, coef * hd.x);
      lds_add(A + 3, coef * hd.y);
      if (wa != 0.0) {
        LdsD *Ba = bacc + 4 * ja;
        lds_add(Ba, -coef * wa * fa * g.x);
        lds_add(Ba + 1, -coef * wa * fa * g.y);
        lds_add(Ba + 2, coef * wa * wa * hd.x);
        lds_add(Ba + 3, coef * wa * wa * hd.y);
      }
      if (wb != 0.0) {
        LdsD *Bb = bacc + 4 * jb;
        lds_add(Bb, -coef * wb * fb * g.x);
        lds_add(Bb + 1, -coef * wb * fb * g.y);
        lds_add(Bb + 2, coef * wb * wb * hd.x);
        lds_add(Bb + 3, coef * wb * wb * hd.y);
      }
      return 0.0;
    }
    if (mode == 1) {
      const double bpp = barrier_bpp(d, dhat);
      const V2<double> dq = ld2(pdir, ip) - wa * ld2(pdir, ia) - wb * ld2(pdir, ib);
      const double dn = dot(dq, n), dt_ = dot(dq, tn);
      const double val = fmax(kappa * (bpp * dn * dn + bp * ddd), 0.0);
      return coef * h2 * (val + kt * dt_ * dt_);
    }
    const double asl = p.epsv / p.inv_dt;
    const double au = fabs(ut);
    const double f0 = au < asl ? au * au / (2.0 * asl) : au - 0.5 * asl;
    return coef * h2 * (kappa * barrier_b(d, dhat) + mu * f_n * f0);
  }

  // Is boundary node jp excluded against boundary node j of this scene (mesh neighbours,
  // nodes of one rigid piece)? From the LDS bitset when it was staged, else from global memory.
  __device__ __forceinline__ bool excluded(int jp, int j) const {
    if (sc.excl_lds) {
      const unsigned idx = (unsigned)jp * (unsigned)nb + (unsigned)j;
      return (xbits[idx >> 5] >> (idx & 31u)) & 1u;
    }
    return bit_of(s.excl_bits + eoff, (int64_t)jp * nb + j);
  }

  // One node-edge pair (reference implicit/ipc_contact.py:175-296, the body of the query loop):
  // edge k (global), scene-local boundary node jp that is neither end nor excluded.
  // What a pair reads from global memory before any arithmetic: the edge's nodes (global and
  // scene-local boundary numbering), its vertex-correction coefficients, node jp's number.
  struct PairIn {
    int k, jp, a, b, la, lb, pn;
    double ca, cb;
  };
  __device__ __forceinline__ PairIn pair_load(int k, int jp) {
    return {k, jp, s.ea[k], s.eb[k], s.ela[k], s.elb[k], lbnd[jp], s.ca[k], s.cb[k]};
  }
  __device__ __forceinline__ void pair_eval(int mode, int k, int jp, GCD *pdir, double t_c, double h2,
                                            double &psum, double &cmin, double &amin) {
    pair_eval(mode, pair_load(k, jp), pdir, t_c, h2, psum, cmin, amin);
  }
  __device__ __forceinline__ void pair_eval(int mode, const PairIn &in, GCD *pdir, double t_c, double h2,
                                            double &psum, double &cmin, double &amin) {
    const double dhat = p.dhat, reach = p.ipc_reach;
    const double dlo = 1.0e-9 * dhat, onset = 0.02 * dhat;
    const int jp = in.jp, a = in.a, b = in.b, la = in.la, lb = in.lb;
    const V2<double> xa = {bpos[2 * la], bpos[2 * la + 1]}, xb = {bpos[2 * lb], bpos[2 * lb + 1]};
    const V2<double> ab = xb - xa;
    const double L2 = fmax(dot(ab, ab), 1.0e-30);
    const double L = sqrt(L2);
    const V2<double> xp = {bpos[2 * jp], bpos[2 * jp + 1]};
    const V2<double> w = xp - xa;
    const double tt = dot(w, ab) / L2;
    V2<double> dvec = w - tt * ab;
    if (tt <= 0.0) dvec = w;
    else if (tt >= 1.0) dvec = xp - xb;
    double d = len(dvec);
    if (!(d < reach)) return;
    const int pn = in.pn;
    if (mode == 1 || mode == 3) {
      cmin = fmin(cmin, d);
      double keep = fmax(p.ipc_gap * d, dhat - onset);
      if (d < dhat) keep = d - fmin(2.0 * (dhat - d) + onset, 0.5 * d);
      if (mode == 3) keep = fmin(d, fmax(p.ipc_gap * d, 0.25 * dhat));
      amin = fmin(amin, accd(xp, xa, xb, ld2(pdir, pn), ld2(pdir, a), ld2(pdir, b), keep, t_c));
      if (mode == 3) return;
    }
    double mu_p = p.ipc_mu;
    if (p.ipc_per_body && s.mu_node) {
      const double mi = s.mu_node[pn], mj = s.mu_node[a], sm = mi + mj;
      mu_p = sm > 0.0 ? 2.0 * mi * mj / sm : 0.0;
    }
    if (tt > 0.0 && tt < 1.0 && d < dhat) {
      d = fmax(d, dlo);
      double ddd = 0.0;
      if (mode == 1) {
        const V2<double> de = ld2(pdir, b) - ld2(pdir, a);
        const V2<double> dw = ld2(pdir, pn) - ld2(pdir, a);
        const double c0_ = cross(ab, w);
        const double c1_ = cross(de, w) + cross(ab, dw);
        const double c2_ = 2.0 * cross(de, dw);
        const double L1 = dot(ab, de) / L;
        const double Lq = (dot(de, de) - L1 * L1) / L;
        const double s0 = c0_ / L;
        const double s1 = (c1_ - s0 * L1) / L;
        ddd = (c2_ - 2.0 * s1 * L1 - s0 * Lq) / L;
        if (c0_ < 0.0) ddd = -ddd;
      }
      psum += term(mode, 1.0, jp, la, lb, pn, a, b, 1.0 - tt, tt, d, (1.0 / d) * dvec, ddd, mu_p, pdir, h2);
    }
    double coef_a = -in.ca;
    if (tt <= 0.0) coef_a += 1.0;
    double d_a = len(w);
    if (coef_a != 0.0 && d_a < dhat) {
      d_a = fmax(d_a, dlo);
      const V2<double> n_a = (1.0 / d_a) * w;
      double ddd = 0.0;
      if (mode == 1) {
        const V2<double> dw = ld2(pdir, pn) - ld2(pdir, a);
        const double dn = dot(dw, n_a);
        ddd = (dot(dw, dw) - dn * dn) / d_a;
      }
      psum += term(mode, coef_a, jp, la, lb, pn, a, b, 1.0, 0.0, d_a, n_a, ddd, mu_p, pdir, h2);
    }
    double coef_b = -in.cb;
    if (tt >= 1.0) coef_b += 1.0;
    const V2<double> wb_ = xp - xb;
    double d_b = len(wb_);
    if (coef_b != 0.0 && d_b < dhat) {
      d_b = fmax(d_b, dlo);
      const V2<double> n_b = (1.0 / d_b) * wb_;
      double ddd = 0.0;
      if (mode == 1) {
        const V2<double> dw = ld2(pdir, pn) - ld2(pdir, b);
        const double dn = dot(dw, n_b);
        ddd = (dot(dw, dw) - dn * dn) / d_b;
      }
      psum += term(mode, coef_b, jp, la, lb, pn, a, b, 0.0, 1.0, d_b, n_b, ddd, mu_p, pdir, h2);
    }
  }

  // Candidates of edge k from the grid: calls f(jp) for every boundary node in the cells the
  // edge's box +- reach overlaps that is neither end of the edge nor excluded against one.
  template <typename F>
  __device__ __forceinline__ void candidates(int k, F &&f) {
    const double reach = p.ipc_reach + p.ipc_skin;
    const int la = s.ela[k], lb = s.elb[k];
    const double ax = bpos[2 * la], ay = bpos[2 * la + 1], bx = bpos[2 * lb], by = bpos[2 * lb + 1];
    int ix0 = (int)((fmin(ax, bx) - reach - sc.gx0) / sc.gh);
    int ix1 = (int)((fmax(ax, bx) + reach - sc.gx0) / sc.gh);
    int iy0 = (int)((fmin(ay, by) - reach - sc.gy0) / sc.gh);
    int iy1 = (int)((fmax(ay, by) + reach - sc.gy0) / sc.gh);
    ix0 = max(ix0, 0);
    iy0 = max(iy0, 0);
    ix1 = min(ix1, sc.gnx - 1);
    iy1 = min(iy1, sc.gny - 1);
    for (int iy = iy0; iy <= iy1; ++iy)
      for (int ix = ix0; ix <= ix1; ++ix) {
        const int c = iy * sc.gnx + ix;
        for (int q = cstart[c]; q < cstart[c + 1]; ++q) {
          const int jp = citem[q];
          if (jp == la || jp == lb || excluded(jp, la) || excluded(jp, lb)) continue;
          f(jp);
        }
      }
  }

  // The pairs within reach at the positions the grid was built from, into the LDS list; the
  // list is used by every pass at these positions unless it overflowed.
  __device__ __forceinline__ void build_pairs() {
    if (tid == 0) sc.npairs = 0;
    __syncthreads();
    const bool fits = s.max_pairs > 0 && k1 - k0 < 32768;  // edge index packed in 15 bits
    if (fits) {
      const double reach = p.ipc_reach + p.ipc_skin;
      for (int k = k0 + tid; k < k1; k += TPB) {
        const int la = s.ela[k], lb = s.elb[k];
        const V2<double> xa = {bpos[2 * la], bpos[2 * la + 1]}, xb = {bpos[2 * lb], bpos[2 * lb + 1]};
        candidates(k, [&](int jp) {
          const V2<double> xp = {bpos[2 * jp], bpos[2 * jp + 1]};
          if (seg_dist(xp, xa, xb) < reach) {
            const int q = lds_add(&sc.npairs, 1);
            if (q < s.max_pairs) plist[q] = ((k - k0) << 16) | jp;
          }
        });
      }
    }
    __syncthreads();
    if (tid == 0) sc.plist = fits && sc.npairs <= s.max_pairs;
    __syncthreads();
  }

  // One pass over the pairs (reference implicit/ipc_contact.py:175-296): mode 0 gradient +
  // diagonal into bacc, 1 p^T H p + clearance + step limit, 2 energy, 3 clearance + step limit.
  __device__ __forceinline__ void ipc_pass(int mode, GCD *pdir, double t_c, double h2, double &psum,
                           double &cmin, double &amin) {
    psum = 0.0;
    cmin = BIG;
    amin = BIG;
    if (sc.plist) {
      // two pairs per thread loaded before either is evaluated (each thread's pairs in the
      // same order as one per trip)
      range_batches<2>(
          0, sc.npairs,
          [&](int q) {
            const int v = plist[q];
            return pair_load(k0 + (v >> 16), v & 0xffff);
          },
          [&](int, const PairIn &in) { pair_eval(mode, in, pdir, t_c, h2, psum, cmin, amin); });
      return;
    }
    for (int k = k0 + tid; k < k1; k += TPB)
      candidates(k, [&](int jp) { pair_eval(mode, k, jp, pdir, t_c, h2, psum, cmin, amin); });
  }

  // ------------------------------------------------------------------ "node" contact
  // Partners of scene-local boundary node j: every other boundary node in the grid cells within
  // node_r + skin of it that is not excluded against j, in cell order (reference
  // implicit/kernels.py:772-869 finds the same set through a hash-grid query). f(jb, xb).
  template <typename F>
  __device__ __forceinline__ void node_partners(int j, V2<double> xa, F &&f) {
    const double reach = p.node_r + p.ipc_skin;
    int ix0 = (int)((xa.x - reach - sc.gx0) / sc.gh), ix1 = (int)((xa.x + reach - sc.gx0) / sc.gh);
    int iy0 = (int)((xa.y - reach - sc.gy0) / sc.gh), iy1 = (int)((xa.y + reach - sc.gy0) / sc.gh);
    ix0 = max(ix0, 0);
    iy0 = max(iy0, 0);
    ix1 = min(ix1, sc.gnx - 1);
    iy1 = min(iy1, sc.gny - 1);
    for (int iy = iy0; iy <= iy1; ++iy)
      for (int ix = ix0; ix <= ix1; ++ix) {
        const int c = iy * sc.gnx + ix;
        for (int q = cstart[c]; q < cstart[c + 1]; ++q) {
          const int jb = citem[q];
          if (jb == j || excluded(j, jb)) continue;
          f(jb, V2<double>{bpos[2 * jb], bpos[2 * jb + 1]});
        }
      }
  }

  // Penalty and pair-friction gradient and diagonal of boundary node j (node n at x), added to
  // g, dg (reference implicit/kernels.py:772-808, implicit/body_contact.py:37-82).
  __device__ __forceinline__ void node_grad(int j, int n, V2<double> x, double h2, V2<double> &g,
                                            V2<double> &dg) {
    V2<double> gp = {0.0, 0.0}, gf = {0.0, 0.0};
    double hx = 0.0, hy = 0.0, fx = 0.0, fy = 0.0;
    const bool fric = p.node_fric;
    const V2<double> ua = fric ? x - ld2(s.xn, n) : V2<double>{0.0, 0.0};
    node_partners(j, x, [&](int jb, V2<double> xb) {
      NodeVisit v;
      if (!node_visit(x, xb, p.node_r, v)) return;
      node_grad_diag(v, p.node_k, 1, gp, hx, hy);
      if (fric) {
        const int nj = lbnd[jb];
        const double mu = node_pair_mu(p.node_mu, p.ipc_per_body, s.mu_node, n, nj);
        node_fric_grad_diag(v, p.node_k, mu, p.epsv, p.inv_dt, ua - (xb - ld2(s.xn, nj)), gf, fx, fy);
      }
    });
    g = g + h2 * gp;
    dg = dg + V2<double>{h2 * hx, h2 * hy};
    if (fric && s.free_[n]) {
      g = g + h2 * gf;
      dg = dg + V2<double>{h2 * fx, h2 * fy};
    }
  }

  // Half of p^T H p (mode 1) or of the energy (mode 2) of boundary node j's pairs, times h2
  // (reference implicit/kernels.py:811-869, implicit/body_contact.py:86-167).
  __device__ __forceinline__ double node_sum(int mode, int j, int n, V2<double> x, double h2) {
    double val = 0.0, valf = 0.0;
    const bool fric = p.node_fric;
    const V2<double> ua = fric ? x - ld2(s.xn, n) : V2<double>{0.0, 0.0};
    const V2<double> pa = mode == 1 ? ld2(s.p, n) : V2<double>{0.0, 0.0};
    node_partners(j, x, [&](int jb, V2<double> xb) {
      NodeVisit v;
      if (!node_visit(x, xb, p.node_r, v)) return;
      const int nj = lbnd[jb];
      const V2<double> dp = mode == 1 ? pa - ld2(s.p, nj) : V2<double>{0.0, 0.0};
      val += mode == 1 ? node_pHp(v, p.node_k, 1, dp) : node_energy(v, p.node_k);
      if (fric) {
        const double mu = node_pair_mu(p.node_mu, p.ipc_per_body, s.mu_node, n, nj);
        const V2<double> du = ua - (xb - ld2(s.xn, nj));
        valf += mode == 1 ? node_fric_pHp(v, p.node_k, mu, p.epsv, p.inv_dt, du, dp)
                          : node_fric_energy(v, p.node_k, mu, p.epsv, p.inv_dt, du);
      }
    });
    return h2 * val + h2 * valf;
  }

  // ------------------------------------------------------------------ kinematic sweep
  // reference implicit/ipc_contact.py:308-331, 456-483
  __device__ __forceinline__ void sweep() {
    double q = 0.0;
    for (int b = bc0 + tid; b < bc1; b += TPB) {
      const int i = s.bc_idx[b];
      const V2<double> d = ld2(s.bc_target, b) - X(i);
      st2(s.bc_q, i, d);
      q = fmax(q, len(d));
    }
    q = blk.max1(q);
    if (tid == 0) sc.q_inf = q;
    contact_prepare();
    double ps, cm, am;
    ipc_pass(3, s.bc_q, 1.0, p.h2, ps, cm, am);
    double v[2] = {cm, am};
    blk.template min<2>(v);
    if (tid == 0) {
      double beta = fmin(1.0, v[1]);
      if (q > 0.0) beta = fmin(beta, 0.45 * p.ipc_reach / q);
      sc.sweep_beta = beta;
      sc.clr_ipc = v[0];
    }
    __syncthreads();
    const double t = sc.sweep_beta;
    for (int b = bc0 + tid; b < bc1; b += TPB) {
      const int i = s.bc_idx[b];
      const V2<double> tg = ld2(s.bc_target, b);
      const V2<double> xi = t >= 1.0 ? tg : X(i) + t * (tg - X(i));
      xput(i, xi);
      st2(s.xhat, i, xi);
    }
    __syncthreads();
  }

  // ------------------------------------------------------------------ gradient + diagonal
  __device__ __forceinline__ void grad_diag() {
    // The gradient buffers change roles each iteration: the one holding the gradient of the
    // iteration before (read by direction() as the previous gradient) is left as it is, and the
    // new gradient is written over the older one, so no copy of it is made.
    GD *const t = gprev;
    gprev = gcur;
    gcur = t;
    if (s.nacc_lds && n1 - n0 <= s.max_nn) {
      grad_diag_lds();
    } else {
      grad_diag_colored();
    }
    grad_diag_rest();
  }

  // Elastic gradient and diagonal (reference implicit/kernels.py:86-128) summed per node in
  // LDS with LDS atomics, all soft triangles in one pass; then one pass over the nodes writes
  // grad = inertia - h^2 f_ext + h^2 elastic, diag = mass + h^2 elastic
  // (reference grad_diag_mass_kernel, implicit/kernels.py:73-82).
  __device__ __forceinline__ void grad_diag_lds() {
    const double h2 = p.h2;
    const int nn = n1 - n0;
    const int ns = s.max_nn;
    for (int j = tid; j < 4 * ns; j += TPB) nacc[j] = 0.0;
    for (int j = tid; j < 4 * nb; j += TPB) bacc[j] = 0.0;
    __syncthreads();
    const int ta = s.sc_cs[c0], tb = s.sc_cs[c0 + MAXC];
    // two triangles per trip, both loaded before either is computed, so that their global
    // loads are in flight together
    const ElemPtrs<R> ep = elem_ptrs<R>(s);
    int t = ta + tid;
    for (; t + TPB < tb; t += 2 * TPB) {
      const Elem<R> e0 = load_elem<R>(ep, t), e1 = load_elem<R>(ep, t + TPB);
      elem_to_lds(e0);
      elem_to_lds(e1);
    }
    if (t < tb) elem_to_lds(load_elem<R>(ep, t));
    __syncthreads();
    struct NodeIn {
      V2<double> x, xh, fe;
      double m, f;
    };
    node_batches<2>(
        [&](int i) {
          NodeIn q;
          q.m = s.mass[i];
          q.f = fr(i);
          q.x = X(i);
          q.xh = ld2(s.xhat, i);
          q.fe = s.fext ? ld2(s.fext, i) : V2<double>{0.0, 0.0};
          return q;
        },
        [&](int i, const NodeIn &q) {
          const V2<double> d = q.x - q.xh;
          const LdsD *a = nacc + (i - n0);
          V2<double> gn = {q.m * d.x * q.f, q.m * d.y * q.f};
          if (s.fext) gn = gn - h2 * q.fe;
          gn = gn + h2 * V2<double>{a[0], a[ns]};
          st2(gcur, i, gn);
          st2(s.diag, i, V2<double>{q.m + h2 * a[2 * ns], q.m + h2 * a[3 * ns]});
        });
    __syncthreads();
  }

  // One triangle's gradient and clamped diagonal into the LDS node accumulators (component
  // k of node l at nacc[k * max_nn + l]: consecutive lanes hit spread-out banks).
  __device__ __forceinline__ void elem_to_lds(const Elem<R> &el) {
    const int ns = s.max_nn;
    const NH<R> h = elem_F(el, X(el.i0), X(el.i1), X(el.i2));
    V2<R> g[3], d[3];
    nh_grad(h, el.B, el.Amu, R(1), g[0], g[1], g[2]);
    nh_diag(h, el.B, el.Amu, el.Alam, R(1), d[0], d[1], d[2]);
    const int li[3] = {el.i0 - n0, el.i1 - n0, el.i2 - n0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      LdsD *a = nacc + li[q];
      lds_add(a, double(g[q].x));
      lds_add(a + ns, double(g[q].y));
      lds_add(a + 2 * ns, fmax(double(d[q].x), 0.0));
      lds_add(a + 3 * ns, fmax(double(d[q].y), 0.0));
    }
  }

  // The same, colour by colour with plain read-modify-write of the node arrays (no two
  // triangles of a colour share a node), for scenes whose nodes do not fit in LDS.
  __device__ __forceinline__ void grad_diag_colored() {
    const double h2 = p.h2;
    struct NodeIn {
      V2<double> x, xh, fe;
      double m, f;
    };
    node_batches<2>(
        [&](int i) {
          NodeIn q;
          q.m = s.mass[i];
          q.f = fr(i);
          q.x = X(i);
          q.xh = ld2(s.xhat, i);
          q.fe = s.fext ? ld2(s.fext, i) : V2<double>{0.0, 0.0};
          return q;
        },
        [&](int i, const NodeIn &q) {
          const V2<double> d = q.x - q.xh;
          V2<double> gn = {q.m * d.x * q.f, q.m * d.y * q.f};
          if (s.fext) gn = gn - h2 * q.fe;
          st2(gcur, i, gn);
          st2(s.diag, i, V2<double>{q.m, q.m});
        });
    for (int j = tid; j < 4 * nb; j += TPB) bacc[j] = 0.0;
    __syncthreads();
    // elastic, colour by colour (reference implicit/kernels.py:86-128). A triangle's three
    // nodes are distinct and no other triangle of its colour touches them, so the old sums of
    // both triangles of a trip are loaded together with their positions before either triangle
    // adds: one memory round trip per trip instead of one per node update.
    struct ColIn {
      Elem<R> el;
      V2<double> x[3], g[3], d[3];
    };
    const ElemPtrs<R> ep = elem_ptrs<R>(s);
    for (int c = 0; c < MAXC; ++c) {
      const int ta = s.sc_cs[c0 + c], tb = s.sc_cs[c0 + c + 1];
      if (ta == tb) continue;
      range_batches<2>(
          ta, tb,
          [&](int t) {
            ColIn q;
            q.el = load_elem<R>(ep, t);
            const int id[3] = {q.el.i0, q.el.i1, q.el.i2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              q.x[k] = X(id[k]);
              q.g[k] = ld2(gcur, id[k]);
              q.d[k] = ld2(s.diag, id[k]);
            }
            return q;
          },
          [&](int, const ColIn &q) {
            const NH<R> h = elem_F(q.el, q.x[0], q.x[1], q.x[2]);
            V2<R> g[3], d[3];
            nh_grad(h, q.el.B, q.el.Amu, R(1), g[0], g[1], g[2]);
            nh_diag(h, q.el.B, q.el.Amu, q.el.Alam, R(1), d[0], d[1], d[2]);
            const int id[3] = {q.el.i0, q.el.i1, q.el.i2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              st2(gcur, id[k], q.g[k] + h2 * up(g[k]));
              st2(s.diag, id[k], q.d[k] + h2 * V2<double>{fmax(double(d[k].x), 0.0), fmax(double(d[k].y), 0.0)});
            }
          });
      __syncthreads();
    }
  }

  // ------------------------------------------------------------------ tendons
  // Tendon k's eyelets are ey_node[ey_start[k] .. ey_start[k + 1]); segment j joins eyelets j and
  // j + 1. The tendon passes take their scratch from the boundary accumulators (free once the
  // boundary pass has read them, until the next gradient pass clears them): per eyelet of the
  // scene its node and tendon, its outgoing segment's vector and length, and its gradient and
  // diagonal terms (8 doubles), and per tendon its tension and stiffness (2).
  __device__ __forceinline__ bool tendon_lds_fits() const {
    const int ne = s.ey_start[td1] - s.ey_start[td0];
    return 8 * ne + 2 * (td1 - td0) <= 4 * s.max_nb;
  }

  // One thread walks the scene's cables in order (the fallback when the scratch does not fit).
  __device__ __forceinline__ void tendons_grad_serial(double h2) {
    for (int k = td0; k < td1; ++k) {
      const int ja = s.ey_start[k], jz = s.ey_start[k + 1];
      double L = 0.0;
      for (int j = ja; j + 1 < jz; ++j) L += len(X(s.ey_node[j + 1]) - X(s.ey_node[j]));
      s.td_L[k] = L;
      double tau, ke;
      if (s.td_stiff[k] > 0.0) {
        const double ext = L - s.td_L0[k];
        tau = ext > 0.0 ? s.td_stiff[k] * ext : 0.0;
        ke = ext > 0.0 ? s.td_stiff[k] : 0.0;
      } else {
        tau = fmax(s.td_T[k], 0.0);
        ke = 0.0;
      }
      s.td_tau[k] = tau;
      s.td_keff[k] = ke;
      for (int j = ja; j < jz; ++j) {
        const int n = s.ey_node[j];
        V2<double> gL = {0.0, 0.0}, geo = {0.0, 0.0};
        const V2<double> xj = X(n);
        if (j > ja) {
          const V2<double> t = xj - X(s.ey_node[j - 1]);
          const double l = fmax(len(t), 1.0e-12), il = 1.0 / l;
          const V2<double> d = {t.x / l, t.y / l};
          gL = gL + d;
          geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
        }
        if (j + 1 < jz) {
          const V2<double> t = X(s.ey_node[j + 1]) - xj;
          const double l = fmax(len(t), 1.0e-12), il = 1.0 / l;
          const V2<double> d = {t.x / l, t.y / l};
          gL = gL - d;
          geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
        }
        st2(gcur, n, ld2(gcur, n) + (h2 * tau) * gL);
        st2(s.diag, n, ld2(s.diag, n) + V2<double>{h2 * (tau * geo.x + ke * gL.x * gL.x),
                                                    h2 * (tau * geo.y + ke * gL.y * gL.y)});
      }
    }
  }

  // The same in parallel, with the same values: every thread takes segments, thread 0 sums
  // each tendon's length in eyelet order and sets its tension, every thread takes eyelets'
  // terms, then the thread of each node's first eyelet adds the node's terms in eyelet order (a
  // node can be an eyelet more than once), so every sum is formed in the order of the walk above.
  __device__ __forceinline__ void tendons_grad_parallel(double h2) {
    const int E0 = s.ey_start[td0], ne = s.ey_start[td1] - E0;
    // per eyelet (index j - E0): node, outgoing segment (x, y, length), gradient and diagonal terms
    LdsD *nd = bacc, *sx = bacc + ne, *sy = bacc + 2 * ne, *sl = bacc + 3 * ne, *cgx = bacc + 4 * ne,
         *cgy = bacc + 5 * ne, *cdx = bacc + 6 * ne, *cdy = bacc + 7 * ne, *tk = bacc + 8 * ne;
    for (int k = td0; k < td1; ++k) {
      const int jz = s.ey_start[k + 1];
      for (int j = s.ey_start[k] + tid; j < jz; j += TPB) {
        const int n = s.ey_node[j];
        nd[j - E0] = double(n);
        if (j + 1 < jz) {
          const V2<double> t = X(s.ey_node[j + 1]) - X(n);
          sx[j - E0] = t.x;
          sy[j - E0] = t.y;
          sl[j - E0] = len(t);
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      for (int k = td0; k < td1; ++k) {
        const int ja = s.ey_start[k], jz = s.ey_start[k + 1];
        double L = 0.0;
        for (int j = ja; j + 1 < jz; ++j) L += sl[j - E0];
        s.td_L[k] = L;
        double tau, ke;
        if (s.td_stiff[k] > 0.0) {
          const double ext = L - s.td_L0[k];
          tau = ext > 0.0 ? s.td_stiff[k] * ext : 0.0;
          ke = ext > 0.0 ? s.td_stiff[k] : 0.0;
        } else {
          tau = fmax(s.td_T[k], 0.0);
          ke = 0.0;
        }
        s.td_tau[k] = tau;
        s.td_keff[k] = ke;
        tk[2 * (k - td0)] = tau;
        tk[2 * (k - td0) + 1] = ke;
      }
    }
    __syncthreads();
    for (int k = td0; k < td1; ++k) {
      const int ja = s.ey_start[k], jz = s.ey_start[k + 1];
      const double tau = tk[2 * (k - td0)], ke = tk[2 * (k - td0) + 1];
      for (int j = ja + tid; j < jz; j += TPB) {
        V2<double> gL = {0.0, 0.0}, geo = {0.0, 0.0};
        if (j > ja) {
          const V2<double> t = {sx[j - 1 - E0], sy[j - 1 - E0]};
          const double l = fmax(sl[j - 1 - E0], 1.0e-12), il = 1.0 / l;
          const V2<double> d = {t.x / l, t.y / l};
          gL = gL + d;
          geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
        }
        if (j + 1 < jz) {
          const V2<double> t = {sx[j - E0], sy[j - E0]};
          const double l = fmax(sl[j - E0], 1.0e-12), il = 1.0 / l;
          const V2<double> d = {t.x / l, t.y / l};
          gL = gL - d;
          geo = geo + V2<double>{(1.0 - d.x * d.x) * il, (1.0 - d.y * d.y) * il};
        }
        const V2<double> cg = (h2 * tau) * gL;
        cgx[j - E0] = cg.x;
        cgy[j - E0] = cg.y;
        cdx[j - E0] = h2 * (tau * geo.x + ke * gL.x * gL.x);
        cdy[j - E0] = h2 * (tau * geo.y + ke * gL.y * gL.y);
      }
    }
    __syncthreads();
    for (int j = tid; j < ne; j += TPB) {
      const double node = nd[j];
      bool first = true;
      for (int i = 0; i < j && first; ++i) first = nd[i] != node;
      if (!first) continue;
      const int n = int(node);
      V2<double> g = ld2(gcur, n), dg = ld2(s.diag, n);
      for (int i = j; i < ne; ++i) {
        if (nd[i] != node) continue;
        g = g + V2<double>{cgx[i], cgy[i]};
        dg = dg + V2<double>{cdx[i], cdy[i]};
      }
      st2(gcur, n, g);
      st2(s.diag, n, dg);
    }
  }

  // The curvature's tendon terms per segment, for thread 0 to sum in order: along = dp . d,
  // w = |dp|^2 - along^2 and 1 / l (3 doubles per eyelet in the same scratch).
  __device__ __forceinline__ void tendon_curv_terms() {
    const int E0 = s.ey_start[td0], ne = s.ey_start[td1] - E0;
    LdsD *sa = bacc, *sw = bacc + ne, *si = bacc + 2 * ne;
    for (int k = td0; k < td1; ++k) {
      const int jz = s.ey_start[k + 1];
      for (int j = s.ey_start[k] + tid; j + 1 < jz; j += TPB) {
        const int na = s.ey_node[j], nbd = s.ey_node[j + 1];
        const V2<double> t = X(nbd) - X(na);
        const double l = fmax(len(t), 1.0e-12), il = 1.0 / l;
        const V2<double> d = {t.x / l, t.y / l};
        const V2<double> dp = ld2(s.p, nbd) - ld2(s.p, na);
        const double along = dot(dp, d);
        sa[j - E0] = along;
        sw[j - E0] = dot(dp, dp) - along * along;
        si[j - E0] = il;
      }
    }
  }

  // Everything after elasticity: rigid-mode diagonal snapshot, contact, pressure, colliders,
  // friction, tendons, affine and rigid-mode gathers.
  __device__ __forceinline__ void grad_diag_rest() {
    const double h2 = p.h2;
    // rigid modes: the diagonal as it stands after elasticity (reference affine.py:541-542)
    if (r1 > r0) {
      for (int g = r0; g < r1; ++g)
        for (int k = s.rn_start[g] + tid; k < s.rn_start[g + 1]; k += TPB) {
          const int i = s.rn_node[k];
          st2(s.diag_el, i, ld2(s.diag, i));
        }
    }
    // contact between bodies and pressure: into the boundary accumulators
    if (p.ipc_on && k1 > k0) {
      contact_prepare();
      double ps, cm, am;
      ipc_pass(0, s.p, 0.0, h2, ps, cm, am);
    }
    for (int k = pe0 + tid; k < pe1; k += TPB) {
      const int a = s.pe_a[k], b = s.pe_b[k];
      const double pr = s.ch_p[s.pe_ch[k]];
      const V2<double> tv = X(b) - X(a);
      const V2<double> g = (-0.5 * h2 * pr) * V2<double>{tv.y, -tv.x};
      const int ja = s.bloc[a], jb = s.bloc[b];
      lds_add(bacc + 4 * ja, g.x);
      lds_add(bacc + 4 * ja + 1, g.y);
      lds_add(bacc + 4 * jb, g.x);
      lds_add(bacc + 4 * jb + 1, g.y);
    }
    __syncthreads();
    if (node_on() && nb > 0) contact_prepare();
    // colliders, friction and the accumulated contact terms, one thread per boundary node
    for (int j = tid; j < nb; j += TPB) {
      const int n = lbnd[j];
      const V2<double> x = X(n);
      V2<double> g = ld2(gcur, n), dg = ld2(s.diag, n);
      for (int k = 0; k < s.n_col; ++k) {
        V2<double> gg, dd;
        if (collider_grad_diag(to_collider_k(s.col[k]), x, p.dhat, p.kappa, h2, 1, gg, dd)) {
          g = g + gg;
          dg = dg + dd;
        }
      }
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider_k(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          g = g + (h2 * c.mu * f.f_n / fmax(f.vtm, p.epsv)) * f.vt;
          const double kt = h2 * c.mu * f.f_n * p.inv_dt / fmax(f.vtm, p.epsv);
          dg = dg + V2<double>{kt * f.n.y * f.n.y, kt * f.n.x * f.n.x};
        }
      }
      if (node_on()) node_grad(j, n, x, h2, g, dg);
      g = g + V2<double>{bacc[4 * j], bacc[4 * j + 1]};
      dg = dg + V2<double>{bacc[4 * j + 2], bacc[4 * j + 3]};
      st2(gcur, n, g);
      st2(s.diag, n, dg);
    }
    __syncthreads();
    // tendons (reference implicit/actuation.py:170-226)
    if (td1 > td0) {
      if (tendon_lds_fits()) tendons_grad_parallel(h2);
      else if (tid == 0) tendons_grad_serial(h2);
      if (sp1 > sp0) __syncthreads();   // the springs below may share nodes with the eyelets
    }
    // springs and pins: one thread, in order (reference implicit/actuation.py:395-413)
    if (sp1 > sp0 && tid == 0) {
      for (int j = sp0; j < sp1; ++j) {
        const int a = s.sp_a[j], b = s.sp_b[j];
        const double k = s.sp_k[j];
        const SpringGeo q = spring_geo(X(a), b >= 0 ? X(b) : ld2(s.sp_anchor, j), k, s.sp_L0[j]);
        const V2<double> g = (h2 * q.tension) * q.n;
        const V2<double> hd = spring_diag(q, k, h2);
        st2(gcur, a, ld2(gcur, a) + g);
        st2(s.diag, a, ld2(s.diag, a) + hd);
        if (b >= 0) {
          st2(gcur, b, ld2(gcur, b) - g);
          st2(s.diag, b, ld2(s.diag, b) + hd);
        }
      }
    }
    if (td1 > td0 || sp1 > sp0) __syncthreads();   // the scene's counts: uniform over the workgroup
    // affine bodies: J^T g, J^T D J per group, then the slaved nodes leave the nodal iteration
    // (reference implicit/affine.py:157-175)
    for (int g = g0; g < g1; ++g) {
      double v[18];
#pragma unroll
      for (int k = 0; k < 18; ++k) v[k] = 0.0;
      for (int k = s.an_start[g] + tid; k < s.an_start[g + 1]; k += TPB) {
        const int i = s.an_node[k];
        const double w0 = s.an_w[3 * k], w1 = s.an_w[3 * k + 1], w2 = s.an_w[3 * k + 2];
        const V2<double> gg = ld2(gcur, i), dd = ld2(s.diag, i);
        st2(gcur, i, V2<double>{0.0, 0.0});   // each slaved node is gathered by one thread
        const double w[3] = {w0, w1, w2};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          v[r] += gg.x * w[r];
          v[3 + r] += gg.y * w[r];
        }
        const double ww[6] = {w0 * w0, w0 * w1, w0 * w2, w1 * w1, w1 * w2, w2 * w2};
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          v[6 + r] += dd.x * ww[r];
          v[12 + r] += dd.y * ww[r];
        }
      }
      double prev[6];                 // last iteration's J^T g, loaded while the sum runs
      if (tid == 0)
        for (int r = 0; r < 6; ++r) prev[r] = s.ag_gq[6 * g + r];
      blk.template sum<18>(v);
      if (tid == 0) {
        for (int r = 0; r < 6; ++r) {
          s.ag_gq_prev[6 * g + r] = prev[r];
          s.ag_gq[6 * g + r] = v[r];
        }
        const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int r = 0; r < 9; ++r) {
          s.ag_Hx[9 * g + r] = v[6 + map[r]];
          s.ag_Hy[9 * g + r] = v[12 + map[r]];
        }
      }
    }
    // rigid-mode coarse space (reference implicit/affine.py:405-453)
    for (int g = r0; g < r1; ++g) {
      double a[3] = {0.0, 0.0, 0.0};
      for (int k = s.rn_start[g] + tid; k < s.rn_start[g + 1]; k += TPB) {
        const int i = s.rn_node[k];
        const double m = s.mass[i];
        const V2<double> x = X(i);
        a[0] += m * x.x;
        a[1] += m * x.y;
        a[2] += m;
      }
      blk.template sum<3>(a);
      const double cx = a[0] / a[2], cy = a[1] / a[2];
      double v[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = 0.0;
      for (int k = s.rn_start[g] + tid; k < s.rn_start[g + 1]; k += TPB) {
        const int i = s.rn_node[k];
        const V2<double> x = X(i);
        const V2<double> rr = {x.x - cx, x.y - cy};
        const V2<double> gg = ld2(gcur, i), gp = ld2(gprev, i);
        const double m = s.mass[i];
        const V2<double> d = ld2(s.diag, i) - ld2(s.diag_el, i);
        const double dx = fmax(d.x, 0.0) + m, dy = fmax(d.y, 0.0) + m;
        v[0] += gg.x;
        v[1] += gg.y;
        v[2] += rr.x * gg.y - rr.y * gg.x;
        v[3] += gp.x;
        v[4] += gp.y;
        v[5] += rr.x * gp.y - rr.y * gp.x;
        v[6] += dx;
        v[7] += -dx * rr.y;
        v[8] += dy;
        v[9] += dy * rr.x;
        v[10] += dx * rr.y * rr.y + dy * rr.x * rr.x;
      }
      blk.template sum<12>(v);
      if (tid == 0) {
        s.rg_acc[3 * g] = a[0];
        s.rg_acc[3 * g + 1] = a[1];
        s.rg_acc[3 * g + 2] = a[2];
        for (int r = 0; r < 3; ++r) {
          s.rg_gq[3 * g + r] = v[r];
          s.rg_gq_prev[3 * g + r] = v[3 + r];
        }
        GD *H = s.rg_H + 9 * g;
        H[0] = v[6]; H[1] = 0.0; H[2] = v[7];
        H[3] = 0.0; H[4] = v[8]; H[5] = v[9];
        H[6] = v[7]; H[7] = v[9]; H[8] = v[10];
      }
    }
    __syncthreads();
  }

  // ------------------------------------------------------------------ Dai-Kou direction
  __device__ __forceinline__ void direction(int it) {
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // With the element accumulators in LDS their planes are free until the next gradient pass:
    // the dot loop parks each node's -P g and f p there, and the update of p reads them back
    // instead of reloading the mask, diagonal, gradient and p from memory. (Lean variant only:
    // in the full one the extra live values spill.)
    const bool park = !FULL && s.nacc_lds && n1 - n0 <= s.max_nn;
    const int ns = s.max_nn;
    struct DirIn {
      V2<double> g, dh, pp, gp;
      double f;
    };
    // The reductions of the affine and rigid-mode groups (a 6x6 Cholesky or a 3x3 inverse
    // each: serial chains of float64 divisions and square roots) read only the groups' own
    // gathered state, not the node arrays this loop writes. With such groups, the last wave
    // does them, one lane per group, while the other waves run the node loop; their
    // contributions to the five dots enter the block sum with the nodes'.
    constexpr int LW = TPB - WAVE;
    static_assert(LW > 0, "the side wave needs at least two waves");
    const bool side = g1 > g0 || r1 > r0;
    if (side && tid >= LW) {
      for (int g = g0 + (tid - LW); g < g1; g += WAVE) affine_reduce(g, v);
      for (int g = r0 + (tid - LW); g < r1; g += WAVE) rigid_reduce(g, v);
    } else {
      range_batches<4>(
        n0, n1, tid, side ? LW : TPB,
        [&](int i) {
          DirIn q;
          q.f = fr(i);
          q.g = ld2(gcur, i);
          q.dh = ld2(s.diag, i);
          q.pp = ld2(s.p, i);
          q.gp = ld2(gprev, i);
          return q;
        },
        [&](int i, const DirIn &q) {
          const double f = q.f;
          const V2<double> g = {q.g.x * f, q.g.y * f};
          st2(gcur, i, g);
          const double Px = f / fmax(q.dh.x, 1.0e-12), Py = f / fmax(q.dh.y, 1.0e-12);
          const V2<double> pp = q.pp;
          const V2<double> yy = g - q.gp;
          const double Pg0 = Px * g.x, Pg1 = Py * g.y, Py0 = Px * yy.x, Py1 = Py * yy.y;
          v[0] += g.x * Pg0 + g.y * Pg1;
          v[1] += g.x * pp.x + g.y * pp.y;
          v[2] += g.x * Py0 + g.y * Py1;
          v[3] += yy.x * pp.x + yy.y * pp.y;
          v[4] += yy.x * Py0 + yy.y * Py1;
          if (park) {
            LdsD *a = nacc + (i - n0);
            a[0] = -Pg0;
            a[ns] = -Pg1;
            a[2 * ns] = pp.x * f;
            a[3 * ns] = pp.y * f;
          }
        });
    }
    blk.template sum<5>(v);
    if (tid == 0) {
      sc.gPg = v[0];
      sc.g_p = v[1];
      sc.g_Py = v[2];
      sc.y_p = v[3];
      sc.y_Py = v[4];
      double b = 0.0;
      if (it > 0) {
        const double ytp = v[3];
        if (fabs(ytp) > 1.0e-30) b = v[2] / ytp - v[4] * v[1] / (ytp * ytp);
      }
      double t = -v[0] + b * v[1];
      if (t >= 0.0) {
        b = 0.0;
        t = -v[0];
      }
      if (p.cg_guard > 0.0 && it > 0) {
        const double hold = (p.ipc_on && sc.pending) ? sc.q_inf : 0.0;
        if (hold <= 0.0 && v[3] < p.cg_guard * sc.alpha * sc.pHp) {
          b = 0.0;
          t = -v[0];
        }
      }
      sc.beta = b;
      sc.gTp = t;
      for (int g = g0; g < g1; ++g)
        for (int r = 0; r < 6; ++r) gpq_(g)[r] = -gzq_(g)[r] + b * gpq_(g)[r];
    }
    __syncthreads();
    const double bb = sc.beta;
    struct UpdIn {
      V2<double> dh, g, pp;
      double f;
    };
    if (park) {
      // p = (-P g + beta p) f, from the parked terms (f is 0 or 1: both terms carry it)
      for (int i = n0 + tid; i < n1; i += TPB) {
        const LdsD *a = nacc + (i - n0);
        st2(s.p, i, V2<double>{a[0] + bb * a[2 * ns], a[ns] + bb * a[3 * ns]});
      }
    } else node_batches<4>(
        [&](int i) {
          UpdIn q;
          q.f = fr(i);
          q.dh = ld2(s.diag, i);
          q.g = ld2(gcur, i);
          q.pp = ld2(s.p, i);
          return q;
        },
        [&](int i, const UpdIn &q) {
          const double f = q.f;
          const double px = -f / fmax(q.dh.x, 1.0e-12) * q.g.x + bb * q.pp.x;
          const double py = -f / fmax(q.dh.y, 1.0e-12) * q.g.y + bb * q.pp.y;
          st2(s.p, i, V2<double>{px * f, py * f});
        });
    __syncthreads();
    for (int g = g0; g < g1; ++g) expand(g, gpq_(g), s.p);
    for (int g = r0; g < r1; ++g) {
      GCD *a = s.rg_acc + 3 * g, *z = s.rg_zq + 3 * g;
      const double cx = a[0] / a[2], cy = a[1] / a[2];
      for (int k = s.rn_start[g] + tid; k < s.rn_start[g + 1]; k += TPB) {
        const int i = s.rn_node[k];
        const V2<double> x = X(i);
        const V2<double> rr = {x.x - cx, x.y - cy};
        st2(s.p, i, ld2(s.p, i) - V2<double>{z[0] - z[2] * rr.y, z[1] + z[2] * rr.x});
      }
    }
    if (g1 > g0 || r1 > r0) __syncthreads();   // without groups nothing was written since the last
  }

  __device__ __forceinline__ void expand(int g, const LdsD *v, GD *out) {
    for (int k = s.an_start[g] + tid; k < s.an_start[g + 1]; k += TPB) {
      GCD *w = s.an_w + 3 * k;
      const V2<double> val = {w[0] * v[0] + w[1] * v[1] + w[2] * v[2], w[0] * v[3] + w[1] * v[4] + w[2] * v[5]};
      if (out == s.x)
        xput(s.an_node[k], val);
      else
        st2(out, s.an_node[k], val);
    }
  }

  // 6x6 block preconditioner of one rigid region (reference implicit/affine.py:179-205)
  __device__ __forceinline__ void affine_reduce(int b, double (&v)[5]) {
    const LdsD *q = gq_(b);
    const double Mu = gMu_(b), La = gLa_(b);
    const NH<double> h = nh_eval(M2<double>{q[1], q[2], q[4], q[5]}, Mu, La);
    const double P[4] = {Mu * h.F.a + h.c * h.G.a, Mu * h.F.b + h.c * h.G.b, Mu * h.F.c + h.c * h.G.c,
                         Mu * h.F.d + h.c * h.G.d};
    const double add[6] = {0.0, P[0], P[1], 0.0, P[2], P[3]};
    double g[6], y[6], z[6], zy[6];
    for (int r = 0; r < 6; ++r) g[r] = s.ag_gq[6 * b + r] + p.h2 * add[r];
    M66 H, L;
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) H.m[r][c] = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        H.m[r][c] = s.ag_Hx[9 * b + 3 * r + c];
        H.m[r + 3][c + 3] = s.ag_Hy[9 * b + 3 * r + c];
      }
    affine_block(H, q[1], q[2], q[4], q[5], Mu, La, p.h2);
    cholesky6(H, L);
    for (int r = 0; r < 6; ++r) y[r] = g[r] - s.ag_gq_prev[6 * b + r];
    cholesky6_solve(L, g, z);
    cholesky6_solve(L, y, zy);
    for (int r = 0; r < 6; ++r) {
      const double pp = gpq_(b)[r];
      v[0] += g[r] * z[r];
      v[1] += g[r] * pp;
      v[2] += g[r] * zy[r];
      v[3] += y[r] * pp;
      v[4] += y[r] * zy[r];
    }
    for (int r = 0; r < 6; ++r) {
      s.ag_gq[6 * b + r] = g[r];
      gzq_(b)[r] = z[r];
    }
  }

  // reference implicit/affine.py:457-473
  __device__ __forceinline__ void rigid_reduce(int b, double (&v)[5]) {
    double Hi[9];
    double Hl[9];
    for (int k = 0; k < 9; ++k) Hl[k] = s.rg_H[9 * b + k];
    invert33(Hl, Hi);
    GCD *g = s.rg_gq + 3 * b, *g0_ = s.rg_gq_prev + 3 * b;
    const double y[3] = {g[0] - g0_[0], g[1] - g0_[1], g[2] - g0_[2]};
    double z[3], zy[3];
    for (int r = 0; r < 3; ++r) {
      z[r] = Hi[3 * r] * g[0] + Hi[3 * r + 1] * g[1] + Hi[3 * r + 2] * g[2];
      zy[r] = Hi[3 * r] * y[0] + Hi[3 * r + 1] * y[1] + Hi[3 * r + 2] * y[2];
    }
    v[0] += g[0] * z[0] + g[1] * z[1] + g[2] * z[2];
    v[2] += g[0] * zy[0] + g[1] * zy[1] + g[2] * zy[2];
    v[4] += y[0] * zy[0] + y[1] * zy[1] + y[2] * zy[2];
    for (int r = 0; r < 3; ++r) s.rg_zq[3 * b + r] = z[r];
  }

  // ------------------------------------------------------------------ curvature and caps
  __device__ __forceinline__ void curvature_caps(int it) {
    const double h2 = p.h2;
    // [0] pHp, [1] p_inf (as -max), [2] alpha min, [3] clr_min
    double ph = 0.0, pinf = 0.0, amin = BIG, cmin = BIG;
    struct MassIn {
      V2<double> pp;
      double m;
    };
    node_batches<4>(
        [&](int i) {
          MassIn q;
          q.pp = ld2(s.p, i);
          q.m = s.mass[i];
          return q;
        },
        [&](int i, const MassIn &q) {
          ph += q.m * dot(q.pp, q.pp);
          pinf = fmax(pinf, len(q.pp));
        });
    // elastic curvature of the soft elements, inversion cap over all elements
    const int tsoft = s.sc_cs[c0 + MAXC];
    if constexpr (sizeof(R) == 4) {
      // Two triangles per trip: both triangles' indices and constants (the constant arrays
      // cover the rigid triangles too) are loaded first, then both direction gathers, then both
      // are evaluated; each thread visits its triangles in the order of one per trip. (The
      // float64 kernel keeps one triangle per trip: there the extra registers spill.)
      const ElemPtrs<R> ep = elem_ptrs<R>(s);
      const auto curv = [&](int t, const Elem<R> &el, V2<double> q0, V2<double> q1, V2<double> q2) {
        const V2<double> x0 = X(el.i0), x1 = X(el.i1), x2 = X(el.i2);
        if (t < tsoft) {
          const NH<R> h = elem_F<R, false>(el, x0, x1, x2);
          const M2<R> dDs = {R(q1.x - q0.x), R(q2.x - q0.x), R(q1.y - q0.y), R(q2.y - q0.y)};
          const M2<R> dF = mat_mul(dDs, el.B);
          ph += h2 * fmax(double(nh_quad(h, dF, el.Amu, el.Alam)), 0.0);
        }
        amin = fmin(amin, inversion_root(x1 - x0, x2 - x0, q1 - q0, q2 - q0, p.area_shrink));
      };
      int t = t0 + tid;
      for (; t + TPB < t1; t += 2 * TPB) {
        const Elem<R> e0 = load_elem<R>(ep, t), e1 = load_elem<R>(ep, t + TPB);
        const V2<double> u0 = ld2(s.p, e0.i0), u1 = ld2(s.p, e0.i1), u2 = ld2(s.p, e0.i2);
        const V2<double> w0 = ld2(s.p, e1.i0), w1 = ld2(s.p, e1.i1), w2 = ld2(s.p, e1.i2);
        curv(t, e0, u0, u1, u2);
        curv(t + TPB, e1, w0, w1, w2);
      }
      if (t < t1) {
        const Elem<R> e0 = load_elem<R>(ep, t);
        curv(t, e0, ld2(s.p, e0.i0), ld2(s.p, e0.i1), ld2(s.p, e0.i2));
      }
    } else {
      for (int t = t0 + tid; t < t1; t += TPB) {
        const int i0 = s.sc_tri[3 * t], i1 = s.sc_tri[3 * t + 1], i2 = s.sc_tri[3 * t + 2];
        const V2<double> x0 = X(i0), x1 = X(i1), x2 = X(i2);
        const V2<double> q0 = ld2(s.p, i0), q1 = ld2(s.p, i1), q2 = ld2(s.p, i2);
        if (t < tsoft) {
          const Elem<R> el = load_elem<R>(s, t);
          const NH<R> h = elem_F<R, false>(el, x0, x1, x2);
          const M2<R> dDs = {R(q1.x - q0.x), R(q2.x - q0.x), R(q1.y - q0.y), R(q2.y - q0.y)};
          const M2<R> dF = mat_mul(dDs, el.B);
          ph += h2 * fmax(double(nh_quad(h, dF, el.Amu, el.Alam)), 0.0);
        }
        amin = fmin(amin, inversion_root(x1 - x0, x2 - x0, q1 - q0, q2 - q0, p.area_shrink));
      }
    }
    // colliders: curvature, clearance, friction curvature, step cap
    for (int j = tid; j < nb; j += TPB) {
      const int n = lbnd[j];
      const V2<double> x = X(n), pp = ld2(s.p, n);
      for (int k = 0; k < s.n_col; ++k) {
        const Collider c = to_collider_k(s.col[k]);
        double dist;
        ph += collider_pHp(c, x, pp, p.dhat, p.kappa, h2, 1, dist);
        cmin = fmin(cmin, dist);
      }
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider_k(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          const double kt = h2 * c.mu * f.f_n * p.inv_dt / fmax(f.vtm, p.epsv);
          const double pt = dot(pp, V2<double>{-f.n.y, f.n.x});
          ph += kt * pt * pt;
        }
      }
      for (int k = 0; k < s.n_col; ++k) amin = fmin(amin, collider_ccd(to_collider_k(s.col[k]), x, pp, p.ccd_eta));
      if (node_on()) ph += node_sum(1, j, n, x, h2);
    }
    // pressure (reference implicit/actuation.py:102-110)
    for (int k = pe0 + tid; k < pe1; k += TPB) {
      const V2<double> pa = ld2(s.p, s.pe_a[k]), pb = ld2(s.p, s.pe_b[k]);
      ph += -h2 * s.ch_p[s.pe_ch[k]] * cross(pa, pb);
    }
    // contact between bodies: p^T H p, clearance and the ACCD step limit from alpha_max
    double ipc_c = BIG;
    if (p.ipc_on && k1 > k0) {
      double ps, cm, am;
      ipc_pass(1, s.p, p.alpha_max, h2, ps, cm, am);
      ph += ps;
      ipc_c = cm;
      amin = fmin(amin, am);
    }
    // the tendons' terms per segment, for thread 0 after the reduction (its barriers make them
    // visible)
    if (td1 > td0 && tendon_lds_fits()) tendon_curv_terms();
    // one reduction: the sum of pHp, then the minima of -p_inf, alpha, clearance, IPC clearance
    double vr[5] = {ph, -pinf, amin, cmin, ipc_c};
    blk.template reduce<1>(vr);
    const double vs[1] = {vr[0]};
    const double vm[4] = {vr[1], vr[2], vr[3], vr[4]};
    if (tid == 0) {
      double pH = vs[0];
      const double p_inf = -vm[0];
      // rigid regions' own curvature and the tendons' (single thread, in order)
      for (int g = g0; g < g1; ++g) {
        const LdsD *q = gq_(g), *pq = gpq_(g);
        const double Mu = gMu_(g), La = gLa_(g);
        const NH<double> h = nh_eval(M2<double>{q[1], q[2], q[4], q[5]}, Mu, La);
        pH += h2 * fmax(nh_quad(h, M2<double>{pq[1], pq[2], pq[4], pq[5]}, Mu, La), 0.0);
      }
      const bool tpar = td1 > td0 && tendon_lds_fits();
      for (int k = td0; k < td1; ++k) {
        double dL = 0.0;
        if (tpar) {
          const int E0 = s.ey_start[td0], ne = s.ey_start[td1] - E0;
          const LdsD *sa = bacc, *sw = bacc + ne, *si = bacc + 2 * ne;
          const double tau = s.td_tau[k];
          for (int j = s.ey_start[k]; j + 1 < s.ey_start[k + 1]; ++j) {
            dL += sa[j - E0];
            pH += h2 * tau * sw[j - E0] * si[j - E0];
          }
        } else {
          for (int j = s.ey_start[k]; j + 1 < s.ey_start[k + 1]; ++j) {
            const int na = s.ey_node[j], nbd = s.ey_node[j + 1];
            const V2<double> t = X(nbd) - X(na);
            const double l = fmax(len(t), 1.0e-12), il = 1.0 / l;
            const V2<double> d = {t.x / l, t.y / l};
            const V2<double> dp = ld2(s.p, nbd) - ld2(s.p, na);
            const double along = dot(dp, d);
            dL += along;
            pH += h2 * s.td_tau[k] * (dot(dp, dp) - along * along) * il;
          }
        }
        pH += h2 * s.td_keff[k] * dL * dL;
      }
      for (int j = sp0; j < sp1; ++j) {
        const int a = s.sp_a[j], b = s.sp_b[j];
        const SpringGeo q = spring_geo(X(a), b >= 0 ? X(b) : ld2(s.sp_anchor, j), s.sp_k[j], s.sp_L0[j]);
        V2<double> dp = ld2(s.p, a);
        if (b >= 0) dp = dp - ld2(s.p, b);
        pH += spring_pHp(q, s.sp_k[j], h2, dp);
      }
      double a = fmin(p.alpha_max, vm[1]);
      if (p.ipc_on && p_inf > 0.0) a = fmin(a, 0.45 * p.ipc_reach / p_inf);
      sc.pHp = pH;
      sc.p_inf = p_inf;
      sc.clr_min = vm[2];
      sc.clr_ipc = vm[3];
      // alpha_finalize (reference implicit/kernels.py:375-423)
      if (sc.active) {
        if (pH > 0.0) a = fmin(a, -sc.gTp / pH);
        if (vm[2] < p.near_factor * p.dhat && p_inf > 0.0) a = fmin(a, 0.5 * p.dhat / p_inf);
        if (a <= 0.0) {
          sc.active = 0;
          a = 0.0;
        } else if (p.tol_x > 0.0 && a * p_inf < p.tol_x) {
          sc.active = 0;
          sc.iters += 1;
        } else {
          const double d = -a * sc.gTp - 0.5 * a * a * pH;
          sc.dE = d;
          if (it == 0) sc.dE0 = fmax(d, 1.0e-300);
          else if (d < p.eps * sc.dE0) sc.active = 0;
          sc.iters += 1;
        }
      }
      sc.alpha = a;
    }
    __syncthreads();
  }

  // ------------------------------------------------------------------ update and line search
  __device__ __forceinline__ void set_x(GCD *from, double a) {
    node_batches<4>([&](int i) { return ld2(from, i) + a * ld2(s.p, i); },
                    [&](int i, const V2<double> &x) { xput(i, x); });
  }

  // line search: q = q0 + a pq, then the slaved nodes from q
  __device__ __forceinline__ void affine_from_q0(double a) {
    if (tid == 0)
      for (int g = g0; g < g1; ++g)
        for (int r = 0; r < 6; ++r) gq_(g)[r] = gq0_(g)[r] + a * gpq_(g)[r];
    __syncthreads();
    for (int g = g0; g < g1; ++g) expand(g, gq_(g), s.x);
    __syncthreads();
  }

  // total incremental potential of the scene (reference implicit/solver.py:453-498)
  __device__ __forceinline__ double energy() {
    const double h2 = p.h2;
    double E = 0.0;
    for (int i = n0 + tid; i < n1; i += TPB) {
      const V2<double> d = X(i) - ld2(s.xhat, i);
      E += 0.5 * s.mass[i] * dot(d, d);
      if (s.fext) E += -h2 * dot(ld2(s.fext, i), X(i) - ld2(s.xn, i));
    }
    const int tsoft = s.sc_cs[c0 + MAXC];
    for (int t = t0 + tid; t < tsoft; t += TPB) {
      const Elem<R> el = load_elem<R>(s, t);
      E += h2 * elem_psi(el, X(el.i0), X(el.i1), X(el.i2));
    }
    if (node_on() && nb > 0) contact_prepare();
    for (int j = tid; j < nb; j += TPB) {
      const int n = lbnd[j];
      const V2<double> x = X(n);
      for (int k = 0; k < s.n_col; ++k) E += collider_energy(to_collider_k(s.col[k]), x, p.dhat, p.kappa, h2);
      if (node_on()) E += node_sum(2, j, n, x, h2);
      if (p.friction && s.free_[n]) {
        const V2<double> xn = ld2(s.xn, n);
        for (int k = 0; k < s.n_col; ++k) {
          const Collider c = to_collider_k(s.col[k]);
          FricFrame f;
          if (!friction_frame(c, x, xn, p.dhat, p.kappa, p.inv_dt, f)) continue;
          const double ut = f.vtm / p.inv_dt, a = p.epsv / p.inv_dt;
          E += h2 * c.mu * f.f_n * (ut < a ? ut * ut / (2.0 * a) : ut - 0.5 * a);
        }
      }
    }
    for (int k = pe0 + tid; k < pe1; k += TPB) {
      const int ch = s.pe_ch[k];
      const V2<double> ref = ld2(s.ch_ref, ch);
      E += -h2 * s.ch_p[ch] * 0.5 * cross(X(s.pe_a[k]) - ref, X(s.pe_b[k]) - ref);
    }
    if (p.ipc_on && k1 > k0) {
      contact_prepare();
      double ps, cm, am;
      ipc_pass(2, s.p, 0.0, h2, ps, cm, am);
      E += ps;
    }
    E = blk.sum1(E);
    if (tid == 0) {
      for (int g = g0; g < g1; ++g) {
        const LdsD *q = gq_(g);
        const NH<double> h = nh_eval(M2<double>{q[1], q[2], q[4], q[5]}, gMu_(g), gLa_(g));
        E += h2 * nh_psi(h, gMu_(g), gLa_(g));
      }
      for (int k = td0; k < td1; ++k) {
        double L = 0.0;
        for (int j = s.ey_start[k]; j + 1 < s.ey_start[k + 1]; ++j)
          L += len(X(s.ey_node[j + 1]) - X(s.ey_node[j]));
        if (s.td_stiff[k] > 0.0) {
          const double ext = L - s.td_L0[k];
          if (ext > 0.0) E += 0.5 * h2 * s.td_stiff[k] * ext * ext;
        } else {
          E += h2 * fmax(s.td_T[k], 0.0) * L;
        }
      }
      for (int j = sp0; j < sp1; ++j) {
        const int b = s.sp_b[j];
        const SpringGeo q = spring_geo(X(s.sp_a[j]), b >= 0 ? X(b) : ld2(s.sp_anchor, j), s.sp_k[j], s.sp_L0[j]);
        const double ext = q.len - s.sp_L0[j];
        E += 0.5 * h2 * s.sp_k[j] * ext * ext;
      }
      sc.Ener = E;
    }
    __syncthreads();
    return sc.Ener;
  }

  __device__ __forceinline__ void update() {
    if (!sc.active) return;  // uniform: sc is shared
    const double a = sc.alpha;
    if (!line_search()) {
      // q += a pq touches only the groups' LDS state, which set_x does not read: thread 0 does
      // it first, and the slaved nodes are expanded from the new q after set_x's barrier.
      if (tid == 0)
        for (int g = g0; g < g1; ++g)
          for (int r = 0; r < 6; ++r) gq_(g)[r] = gq_(g)[r] + a * gpq_(g)[r];
      set_x(s.x, a);
      __syncthreads();
      if (g1 > g0) {
        for (int g = g0; g < g1; ++g) expand(g, gq_(g), s.x);
        __syncthreads();
      }
      return;
    }
    // energy backtracking gate (reference implicit/solver.py:403-425)
    const double E0 = energy();
    for (int i = n0 + tid; i < n1; i += TPB) st2(s.x0, i, X(i));
    if (tid == 0)
      for (int g = g0; g < g1; ++g)
        for (int r = 0; r < 6; ++r) gq0_(g)[r] = gq_(g)[r];
    __syncthreads();
    for (int bt = 0; bt < p.n_backtrack; ++bt) {
      set_x(s.x0, sc.alpha);
      __syncthreads();
      if (g1 > g0) affine_from_q0(sc.alpha);
      const double E1 = energy();
      if (tid == 0 && E1 > E0 + p.ls_tol * (fabs(E0) + 1.0)) sc.alpha *= 0.5;
      __syncthreads();
    }
    set_x(s.x0, sc.alpha);
    __syncthreads();
    if (g1 > g0) affine_from_q0(sc.alpha);
  }

  // ------------------------------------------------------------------ the step
  __device__ __forceinline__ void run() {
    // exclusion bitset of this scene into LDS when it fits
    if (tid == 0) {
      sc.excl_lds = 0;
      sc.stale = 0;
      sc.plist = 0;
    }
    __syncthreads();
    if (((p.ipc_on && k1 > k0) || node_on()) && s.excl_bits) {
      const int64_t w0 = s.excl_bits_off[e], nw = s.excl_bits_off[e + 1] - w0;
      if (nw <= s.excl_bits_lds) {
        for (int64_t i = tid; i < nw; i += TPB) xbits[i] = s.excl_bits[w0 + i];
        if (tid == 0) sc.excl_lds = 1;
      }
    }
    __syncthreads();
    for (int i = n0 + tid; i < n1; i += TPB) {
      const V2<double> x = ld2(s.x, i);
      if (xcon) {
        xc[2 * (i - n0)] = x.x;
        xc[2 * (i - n0) + 1] = x.y;
      }
      st2(s.xn, i, x);
      st2(s.xhat, i, x + (p.dt * p.drag) * ld2(s.v, i));
    }
    for (int j = tid; j < nb; j += TPB) lbnd[j] = s.bnd[b0 + j];
    for (int j = tid; j < GST * (g1 - g0); j += TPB) {
      const int g = g0 + j / GST, r = j % GST;
      double v = 0.0;
      if (r < 6) v = s.ag_q[6 * g + r];
      else if (r >= 12 && r < 18) v = s.ag_pq[6 * g + r - 12];
      else if (r >= 18 && r < 24) v = s.ag_zq[6 * g + r - 18];
      else if (r == 24) v = s.ag_Mu[g];
      else if (r == 25) v = s.ag_La[g];
      gst[j] = v;
    }
    if (tid == 0) {
      sc.pending = 0;
      sc.q_inf = 0.0;
      sc.sweep_beta = 1.0;
      sc.alpha = 0.0;
      sc.pHp = 0.0;
    }
    __syncthreads();
    if (!p.ipc_on) {
      for (int b = bc0 + tid; b < bc1; b += TPB) {
        const int n = s.bc_idx[b];
        const V2<double> t = ld2(s.bc_target, b);
        xput(n, t);
        st2(s.xhat, n, t);
      }
    } else if (bc1 > bc0) {
      sweep();
      if (tid == 0) sc.pending = !(sc.sweep_beta >= 1.0);
    }
    if (tid == 0) {
      sc.dE0 = 0.0;
      sc.active = 1;
      sc.iters = 0;
    }
    for (int i = n0 + tid; i < n1; i += TPB) {
      st2(s.p, i, V2<double>{0.0, 0.0});
      st2(gcur, i, V2<double>{0.0, 0.0});
    }
    __syncthreads();
    for (int it = 0; it < p.iter_max; ++it) {
      if (p.ipc_on && sc.pending) {
        sweep();
        if (tid == 0) {
          if (sc.q_inf > 0.0) sc.active = 1;
          if (it % 8 == 7) sc.pending = !(sc.sweep_beta >= 1.0);
        }
        __syncthreads();
      }
      if (p.conv_check && it && it % p.conv_check == 0 && !sc.active) break;
      grad_diag();
      direction(it);
      curvature_caps(it);
      update();   // ends on a barrier (or returns right after curvature_caps's last one)
    }
    for (int i = n0 + tid; i < n1; i += TPB) st2(s.v, i, p.inv_dt * (X(i) - ld2(s.xn, i)));
    for (int j = tid; j < 18 * (g1 - g0); j += TPB) {
      const int g = g0 + j / 18, r = j % 18;
      if (r < 6) s.ag_q[6 * g + r] = gq_(g)[r];
      else if (r < 12) s.ag_pq[6 * g + r - 6] = gpq_(g)[r - 6];
      else s.ag_zq[6 * g + r - 12] = gzq_(g)[r - 12];
    }
    if (tid == 0) {
      s.active[e] = sc.active;
      s.iters[e] = sc.iters;
      s.alpha[e] = sc.alpha;
      s.pHp[e] = sc.pHp;
      s.p_inf[e] = sc.p_inf;
      s.clr_min[e] = sc.clr_min;
      s.clr_ipc[e] = sc.clr_ipc;
      s.dE[e] = sc.dE;
      s.dE0[e] = sc.dE0;
      s.beta[e] = sc.beta;
      s.gTp[e] = sc.gTp;
      s.gPg[e] = sc.gPg;
      s.g_p[e] = sc.g_p;
      s.g_Py[e] = sc.g_Py;
      s.y_p[e] = sc.y_p;
      s.y_Py[e] = sc.y_Py;
      s.q_inf[e] = sc.q_inf;
      s.sweep_beta[e] = sc.sweep_beta;
    }
  }
};

// offset of the second kernel argument (the params) in the kernel-argument segment
constexpr size_t PM_OFFSET = (sizeof(SH) + alignof(Pm) - 1) / alignof(Pm) * alignof(Pm);
static_assert(alignof(SH) == 8 && alignof(Pm) == 8, "kernel-argument layout");

template <typename R, int TPB, bool FULL>
__global__ __launch_bounds__(TPB, TPB == 256 ? 2 : 1) void k_scene(SH s, Pm p) {
  extern __shared__ double lds[];
  __shared__ SceneScalars sc;
  // The arguments are read in place from the kernel-argument segment (s at offset 0, p after
  // it): a reference to a by-value argument would make the compiler copy it into every lane's
  // private memory, and a copy in LDS would hide that its pointers are global.
  S *ks = (S *)__builtin_amdgcn_kernarg_segment_ptr();
  KPm *kp = (KPm *)((const __attribute__((address_space(4))) char *)ks + PM_OFFSET);
  Step<R, TPB, FULL> st(*ks, *kp, *(LdsScalars *)&sc, (LdsD *)lds);
  st.run();
}

template <typename R, int TPB, bool FULL>
static int launch_v(const SH &s, const Pm &p, hipStream_t stream) {
  const size_t lds = sizeof(double) * (Block<TPB>::RED_SIZE + 8 * (size_t)s.max_nb +
                                       Step<R, TPB, FULL>::GST * (size_t)s.max_ag +
                                       (s.nacc_lds ? 4 * (size_t)s.max_nn : 0) +
                                       (s.xc_lds ? 2 * (size_t)s.max_nn : 0)) +
                     sizeof(int) * (2 * (size_t)s.max_cells + 1 + 2 * (size_t)s.max_nb + (size_t)s.excl_bits_lds +
                                    (size_t)s.max_pairs);
  if (lds > 150 * 1024) {
    set_error("d2f_step_scene: %zu bytes of LDS per scene (max_nb=%d, max_cells=%d) exceed the 150 KiB "
              "budget; use the per-kernel path for scenes this large", lds, s.max_nb, s.max_cells);
    return D2F_ECAP;
  }
  if (lds > 64 * 1024) {
    // once per device and size, so that a launch under graph capture makes no attribute call
    static std::atomic<int> lds_set[64];
    int dev = 0;
    D2F_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || lds_set[dev].load() < (int)lds) {
      D2F_HIP(hipFuncSetAttribute((const void *)k_scene<R, TPB, FULL>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      if (dev >= 0 && dev < 64) lds_set[dev].store((int)lds);
    }
  }
  const int n_wg = s.sc_list ? s.n_list : s.B;
  if (n_wg == 0) return 0;
  hipLaunchKernelGGL((k_scene<R, TPB, FULL>), dim3(n_wg), dim3(TPB), lds, stream, s, p);
  D2F_HIP(hipGetLastError());
  return 0;
}

template <typename R, int TPB>
static int launch(const SH &s, const Pm &p, hipStream_t stream) {
  const bool full = s.n_td > 0 || s.n_sp > 0 || s.n_pe > 0 || s.n_rg > 0 || p.node_on || p.line_search;
  return full ? launch_v<R, TPB, true>(s, p, stream) : launch_v<R, TPB, false>(s, p, stream);
}

// The float and float64 kernels are compiled as two objects of this file (D2F_SCENE_PART 0 and
// 1), so that the build compiles them in parallel.
#ifndef D2F_SCENE_PART
#define D2F_SCENE_PART 0
#endif
int scene_launch_f32(const SH &s, const Pm &p, int tpb, hipStream_t st);
int scene_launch_f64(const SH &s, const Pm &p, int tpb, hipStream_t st);
#if D2F_SCENE_PART == 0
int scene_launch_f32(const SH &s, const Pm &p, int tpb, hipStream_t st) {
  if (tpb == 512) return launch<float, 512>(s, p, st);
  if (tpb == 1024) return launch<float, 1024>(s, p, st);
  return launch<float, 256>(s, p, st);
}
#else
int scene_launch_f64(const SH &s, const Pm &p, int tpb, hipStream_t st) {
  if (tpb == 512) return launch<double, 512>(s, p, st);
  if (tpb == 1024) return launch<double, 1024>(s, p, st);
  return launch<double, 256>(s, p, st);
}
#endif

}  // namespace d2f

#if D2F_SCENE_PART == 0
using namespace d2f;

extern "C" int d2f_step_scene(const d2f_scene *s, const d2f_params *p, void *stream) {
  D2F_CHECK(s && p, "d2f_step_scene: null scene or params");
  D2F_CHECK(s->B > 0 && s->N > 0 && s->sc_tri && s->sc_cs && s->sc_tri_start,
            "d2f_step_scene: the scene has no per-scene layout");
  D2F_CHECK(!s->sc_list || (s->n_list >= 0 && s->n_list <= s->B), "d2f_step_scene: bad scene list");
  D2F_CHECK(s->max_cells >= 1 && s->max_nb >= 0, "d2f_step_scene: bad grid capacity");
  D2F_CHECK(s->excl_bits_lds >= 0 && s->max_pairs >= 0, "d2f_step_scene: negative LDS capacity");
  D2F_CHECK(s->n_ag == 0 || (s->max_ag > 0 && s->ag_start),
            "d2f_step_scene: affine groups need ag_start and max_ag >= every scene's group count");
  D2F_CHECK(s->excl_bits_lds == 0 || (s->excl_bits && s->excl_bits_off),
            "d2f_step_scene: excl_bits_lds set without the packed exclusion bitset");
  D2F_CHECK(s->max_pairs == 0 || s->max_nb < 65536,
            "d2f_step_scene: the pair list packs boundary indices in 16 bits (max_nb < 65536)");
  D2F_CHECK(!p->ipc_on || (s->bloc && s->e_start), "d2f_step_scene: IPC needs bloc and e_start");
  D2F_CHECK(!p->node_on || (!p->ipc_on && s->excl_bits && s->excl_bits_off && p->node_r > 0.0 && p->ipc_skin >= 0.0),
            "d2f_step_scene: node contact needs excl_bits, excl_bits_off, node_r > 0 and no IPC contact");
  hipStream_t st = (hipStream_t)stream;
  const int tpb = p->block_size > 0 ? p->block_size : 256;
  if (p->precision == 1) {
    D2F_CHECK(s->sc_Bm32 && s->sc_Amu32 && s->sc_Alam32, "d2f_step_scene: mixed precision needs float32 constants");
    return scene_launch_f32(*s, *p, tpb, st);
  }
  D2F_CHECK(s->sc_Bm && s->sc_Amu && s->sc_Alam, "d2f_step_scene: float64 constants missing");
  return scene_launch_f64(*s, *p, tpb, st);
}

// ---- d2f_step_scene_noalloc: the scene kernel's node work arrays in a static device array
namespace d2f {
// xn, xhat, grad, grad_prev, p, diag, diag_el, x0, bc_q: 2 N doubles each, each array starting
// on a 256-byte boundary. Every device holds its own copy until the process ends.
constexpr int NOALLOC_ARRAYS = 9;
static __device__ __align__(256) double scratch[NOALLOC_ARRAYS * 2 * (size_t)D2F_NOALLOC_MAX_NODES];

__global__ void k_fill_zero(double *a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

constexpr int NOALLOC_DEVICES = 64;
static std::mutex scratch_mu;
static double *scratch_dev[NOALLOC_DEVICES];   // scratch's address on device d, from its first call

// The current device's address of `scratch`, looked up once per device with that device current.
static int scratch_of_current_device(double **out) {
  int dev = 0;
  D2F_HIP(hipGetDevice(&dev));
  D2F_CHECK(dev >= 0 && dev < NOALLOC_DEVICES, "d2f_step_scene_noalloc: device %d beyond %d", dev,
            NOALLOC_DEVICES);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *ptr = nullptr;
    D2F_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    scratch_dev[dev] = (double *)ptr;
  }
  *out = scratch_dev[dev];
  return 0;
}
}  // namespace d2f

extern "C" int d2f_step_scene_noalloc(const d2f_scene *s, const d2f_params *p, void *stream) {
  D2F_CHECK(s && p, "d2f_step_scene_noalloc: null scene or params");
  D2F_CHECK(s->N > 0 && s->N <= D2F_NOALLOC_MAX_NODES,
            "d2f_step_scene_noalloc: %d nodes; the static work arrays hold at most %d "
            "(use d2f_step_scene with caller-owned work arrays)", s->N, D2F_NOALLOC_MAX_NODES);
  double *w = nullptr;
  if (const int rc = scratch_of_current_device(&w)) return rc;
  const size_t n2 = (2 * (size_t)s->N + 31) & ~(size_t)31;   // 256-byte aligned stride
  d2f_scene t = *s;
  double **arr[NOALLOC_ARRAYS] = {&t.xn, &t.xhat, &t.grad, &t.grad_prev, &t.p,
                                  &t.diag, &t.diag_el, &t.x0, &t.bc_q};
  for (int k = 0; k < NOALLOC_ARRAYS; ++k) *arr[k] = w + k * n2;
  // the sweep direction must be zero off the kinematic nodes, whatever the last call left there;
  // the kernel writes every other array before it reads it
  hipStream_t st = (hipStream_t)stream;
  if (s->n_bc > 0) {
    const int64_t n = 2 * (int64_t)s->N;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0,
                       st, t.bc_q, n);
    D2F_HIP(hipGetLastError());
  }
  return d2f_step_scene(&t, p, stream);
}
#endif
