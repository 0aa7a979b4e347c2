// Reverse-mode (adjoint) building blocks of the gradient path.
//
// The reference differentiates evolve_perturbations with jax.jacfwd
// (nb_minimal_example.ipynb:297) straight through diffrax's loop and
// Rodas5Transformed.step (ode_integrators_stiff.py:718-834).  Here the same
// derivative of the discrete solution is accumulated backwards, one
// wavefront per mode: the accepted states come from the checkpoint tape,
// each step is recomputed and then reversed stage by stage.  The step-size
// controller's decisions carry no derivative; every accepted step keeps its
// fraction of [t0, t1] while t0 (start time), t1 (last output time), the
// initial state, the cosmology scalars and every spline knot move.
//
// Parameter derivatives are projected on up to MAXDIR input tangent
// directions (d cosmo, d tables, d tau_out per direction, supplied by the
// caller): every contraction below runs in dn<1+MAXDIR> arithmetic whose
// component 0 is the derivative along the scale factor a (or tau) and whose
// components 1..MAXDIR are the input tangents.
#pragma once

#include "deb_forward.h"

namespace deb {

constexpr int MAXDIR = DEB_MAXDIR;
using TD = dn<1 + MAXDIR>;          // (a or tau, theta_1..theta_MAXDIR)

// The row passes of the reverse step are real calls rather than inlined:
// inlined into the step they leave one kernel body with thousands of spilled
// registers; as calls each is register-allocated on its own.  An outlined
// function must stay small enough to need no long branches: the compiler's
// long-branch sequences in a non-kernel function were seen to use the return
// address registers s[30:31] without saving them (a wild return).  The
// device-assembly test of the build checks this for every outlined function.
#define DEB_OUTLINE __device__ __noinline__

// Rodas5 tableau in flattened form (constant memory, read with wave-uniform
// indices): stage point u_i = y0 + sum_j kA[i][j] k_j, y1 = y0 + sum_j kA8[j] k_j,
// stage time t0 + kc[i] h, kd[i] h dT added in stage i, kC[i][j] / h k_j.
__constant__ double kA[8][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {R5::a21, 0, 0, 0, 0, 0, 0, 0},
    {R5::a31, R5::a32, 0, 0, 0, 0, 0, 0},
    {R5::a41, R5::a42, R5::a43, 0, 0, 0, 0, 0},
    {R5::a51, R5::a52, R5::a53, R5::a54, 0, 0, 0, 0},
    {R5::a61, R5::a62, R5::a63, R5::a64, R5::a65, 0, 0, 0},
    {R5::a61, R5::a62, R5::a63, R5::a64, R5::a65, 1.0, 0, 0},
    {R5::a61, R5::a62, R5::a63, R5::a64, R5::a65, 1.0, 1.0, 0}};
__constant__ double kA8[8] = {R5::a61, R5::a62, R5::a63, R5::a64, R5::a65, 1.0, 1.0, 1.0};
__constant__ double kC[8][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {R5::C21, 0, 0, 0, 0, 0, 0, 0},
    {R5::C31, R5::C32, 0, 0, 0, 0, 0, 0},
    {R5::C41, R5::C42, R5::C43, 0, 0, 0, 0, 0},
    {R5::C51, R5::C52, R5::C53, R5::C54, 0, 0, 0, 0},
    {R5::C61, R5::C62, R5::C63, R5::C64, R5::C65, 0, 0, 0},
    {R5::C71, R5::C72, R5::C73, R5::C74, R5::C75, R5::C76, 0, 0},
    {R5::C81, R5::C82, R5::C83, R5::C84, R5::C85, R5::C86, R5::C87, 0}};
__constant__ double kc[8] = {0.0, R5::c2, R5::c3, R5::c4, R5::c5, 1.0, 1.0, 1.0};
__constant__ double kd[8] = {R5::d1, R5::d2, R5::d3, R5::d4, R5::d5, 0.0, 0.0, 0.0};

// ---------------------------------------------------------------------------
// cross-lane moves and reductions of composite scalars, component by
// component (member access only)
template <class T>
struct ncomp;
template <>
struct ncomp<double> { static constexpr int value = 1; };
template <>
struct ncomp<dual> { static constexpr int value = 2; };
template <int N, class B>
struct ncomp<dn<N, B>> { static constexpr int value = (N + 1) * ncomp<B>::value; };

__device__ __forceinline__ double &comp(double &x, int) { return x; }
__device__ __forceinline__ double &comp(dual &x, int i) { return i == 0 ? x.v : x.d; }
template <int N, class B>
__device__ __forceinline__ double &comp(dn<N, B> &x, int i) {
  constexpr int cb = ncomp<B>::value;
  return i < cb ? comp(x.v, i) : comp(x.d[(i - cb) / cb], (i - cb) % cb);
}

template <class T>
__device__ __forceinline__ T t_from_prev(T x) {
#pragma unroll
  for (int i = 0; i < ncomp<T>::value; ++i) comp(x, i) = from_prev(comp(x, i));
  return x;
}
template <class T>
__device__ __forceinline__ T t_from_next(T x) {
#pragma unroll
  for (int i = 0; i < ncomp<T>::value; ++i) comp(x, i) = from_next(comp(x, i));
  return x;
}
template <class T>
__device__ __forceinline__ T t_readlane(T x, int lane) {
#pragma unroll
  for (int i = 0; i < ncomp<T>::value; ++i) comp(x, i) = readlane(comp(x, i), lane);
  return x;
}
// wave-uniform sums of NV composite values
template <class T, int NV>
__device__ __forceinline__ void t_wave_sum(T (&v)[NV]) {
  constexpr int C = ncomp<T>::value;
  double flat[NV * C];
#pragma unroll
  for (int j = 0; j < NV; ++j)
#pragma unroll
    for (int i = 0; i < C; ++i) flat[j * C + i] = comp(v[j], i);
  wave_sum<NV * C>(flat);
#pragma unroll
  for (int j = 0; j < NV; ++j)
#pragma unroll
    for (int i = 0; i < C; ++i) comp(v[j], i) = flat[j * C + i];
}

// ---------------------------------------------------------------------------
// transposed structured solves: the forward solves' operations in reverse
// order, each replaced by its transpose (same factors, no new factorisation)

// x <- W_B^{-T} x
template <int NS>
__device__ __forceinline__ void bsolveT(const Mode<NS> &m, double (&x)[NS]) {
  const int lane = lane_id();
  // leaf back-substitution x_dg = (r_dg - wdg x_tg) ddg
  const double xdg = readlane(x[0], L_DG);
  double rdgb = m.u->ddg * xdg;
  if (lane == L_TG) x[0] -= m.u->wdg * m.u->ddg * xdg;
  if (lane == L_DG) x[0] = 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) x[j] *= m.binv[j];
  // PCR levels in reverse: x_i <- sg_i x_i + al_i x_{i-s} + be_i x_{i+s}
#pragma unroll
  for (int L = PCR_LEVELS - 1; L >= 0; --L) {
    if (L >= m.nlev) continue;
    const int s = 1 << L;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double v = m.al(L, j) * x[j];
      const double w = m.be(L, j) * x[j];
      double vn, wp;
      if (L == 0) {
        vn = from_next(v);
        wp = from_prev(w);
      } else {
        const int lp = lane + s, lm = lane - s;
        vn = from_lane(v, lp > WAVE - 1 ? WAVE - 1 : lp);
        wp = from_lane(w, lm < 0 ? 0 : lm);
        if (lp > WAVE - 1) vn = 0.0;
        if (lm < 0) wp = 0.0;
      }
      const double sg = (L == 0) ? m.sig0[j] : 1.0;
      x[j] = sg * x[j] + vn + wp;
    }
  }
  // leaf pre-elimination x_tg -= leafm r_dg
  rdgb -= m.u->leafm * readlane(x[0], L_TG);
  if (lane == L_DG) x[0] += rdgb;
}

// x <- W^{-T} x (transpose of wsolve); a call, so that the step factors stay
// in memory rather than in registers across the reverse stage loop
template <int NS>
DEB_OUTLINE void wsolveT(const Mode<NS> &m, double (&xio)[NS]) {
  const int lane = lane_id();
  // a private copy (stores through xio could alias m for the compiler)
  double x[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) x[s] = xio[s];
  const double kba = readlane(x[0], L_A), kb1 = readlane(x[0], L_AHP);
  if (lane == L_A || lane == L_AHP) x[0] = 0.0;
  const double r1b = m.u->gh * kb1;
  double kab = kba + m.u->gh * m.u->J1a * kb1;
  const double c1kb = m.u->gh * kb1;
  double part[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    part[0] += m.Z[0][s] * x[s];
    part[1] += m.Z[1][s] * x[s];
    part[2] += m.Z[2][s] * x[s];
  }
  wave_sum<3>(part);
  double tt[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) tt[r] = m.u->c1Z[r] * c1kb + part[r];
  double gb[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) gb[c] = m.u->Sinv[0][c] * tt[0] + m.u->Sinv[1][c] * tt[1] + m.u->Sinv[2][c] * tt[2];
#pragma unroll
  for (int s = 0; s < NS; ++s)
    x[s] += m.Cf[0][s] * gb[0] + m.Cf[1][s] * gb[1] + m.Cf[2][s] * gb[2] + m.c1[s] * c1kb;
  bsolveT<NS>(m, x);
  double pj = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) pj += m.jcol[s] * x[s];
  kab += wave_sum1(pj);
  if (lane == L_A) x[0] = m.u->inv_kaden * kab;
  if (lane == L_AHP) x[0] = r1b;
#pragma unroll
  for (int s = 0; s < NS; ++s) xio[s] = x[s];
}

// ---------------------------------------------------------------------------
// right-hand side contractions

// ---------------------------------------------------------------------------
// First-order contraction through background features.
//
// Every coefficient of row_visit depends on (a, theta) only through a few
// wave-uniform features (aH, 1/aH, opac, cs2, pb43, the density weights
// grhom Omega_c / a, ..., w, c_a^2, cs2_DE, a itself, a m_nu) and, for
// massive-nu rows, the bin velocity v_q = (1 + (a amnu / q)^2)^(-1/2), whose
// partials are folded into those of a and amnu.  The derivative of
// sum_rows sum_fields mul * coef is therefore sum_m Fbar_m dF_m, where Fbar_m
// (plain doubles) are the partials of the row formulas below and dF_m the
// features' tangents.
enum : int {
  FT_A = 0, FT_AH, FT_IAH, FT_OPAC, FT_CS2, FT_PB43, FT_RC, FT_RB, FT_RG, FT_RR, FT_RN, FT_RQ, FT_W, FT_CA2,
  FT_CS2Q, FT_AMNU, FT_N
};

template <class T>
__device__ __forceinline__ void features_of(const Ctx &cx, T a, const Bg<T> &b, T (&F)[FT_N]) {
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), OmegaDE = cv<T>(cx, F_OMEGADE);
  const P Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU);
  const T a2 = a * a;
  F[FT_A] = a;
  F[FT_AH] = b.aH;
  F[FT_IAH] = 1.0 / b.aH;
  F[FT_OPAC] = b.opac;
  F[FT_CS2] = b.cs2;
  F[FT_PB43] = b.pb43;
  F[FT_RC] = grhom * (Omegam - Omegab) / a;
  F[FT_RB] = grhom * Omegab / a;
  F[FT_RG] = grhog / a2;
  F[FT_RR] = grhor * Neff / a2;
  F[FT_RN] = grhor * Nmnu / a2;
  F[FT_RQ] = grhom * OmegaDE * b.rhoQ * a2;
  F[FT_W] = b.w;
  F[FT_CA2] = b.ca2;
  F[FT_CS2Q] = tconst<T>(0.0) + cv<T>(cx, F_CS2DE);
  F[FT_AMNU] = tconst<T>(0.0) + cv<T>(cx, F_AMNU);
}

template <class T>
__device__ __forceinline__ void features_t(Ctx &cx, T a, T (&F)[FT_N]) {
  const Bg<T> b = background<T>(cx, a);
  features_of<T>(cx, a, b, F);
}

// Values of the functionals of Y and of the "where they enter" sums of w that
// the first-order contraction below needs, plus (optionally) the transposed
// right-hand side A(a)^T w -- one plain-double pass over the rows.
template <class T>
struct RowSumsT {
  T H, E, Pi, F1;  // sum c_{h,eta,pi,1} Y
  T sh, se, sp;    // sum u_{h,eta,pi} w
};
using RowSums = RowSumsT<double>;

// With DFA, also the features' derivatives along a (dfa_out), from a dual
// evaluation beside the plain one (the plain values feed the rows unchanged).
template <int NS, bool TAPPLY, bool DFA = false>
__device__ __forceinline__ RowSums coef_pass_body(Ctx &cx, const uint32_t (&rd)[NS], double a, double invtau,
                                                  const double (&w_in)[NS], const double (&Y_in)[NS],
                                                  double (&out)[NS], double (&fv_out)[FT_N],
                                                  double *dfa_out = nullptr) {
  const int lane = lane_id();
  double w[NS], Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    w[s] = w_in[s];
    Y[s] = Y_in[s];
  }
  const Bg<double> b = background<double>(cx, a);
  double fv[FT_N];
  features_of<double>(cx, a, b, fv);
#pragma unroll
  for (int m = 0; m < FT_N; ++m) fv_out[m] = fv[m];
  if constexpr (DFA) {
    dual ad;
    ad.v = a;
    ad.d = 1.0;
    dual Fd[FT_N];
    features_t<dual>(cx, ad, Fd);
#pragma unroll
    for (int m = 0; m < FT_N; ++m) dfa_out[m] = Fd[m].d;
  }
  double o[NS];
  double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double cc[4][NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Coef<double> r = row_coef_slot<double>(rd[s], b, cx, invtau, s == 0);
    part[0] += r.ch * Y[s];
    part[1] += r.ceta * Y[s];
    part[2] += r.cpi * Y[s];
    part[3] += r.c1 * Y[s];
    part[4] += r.uh * w[s];
    part[5] += r.ueta * w[s];
    part[6] += r.upi * w[s];
    if constexpr (TAPPLY) {
      o[s] = r.diag * w[s] + from_next(r.sub * w[s]) + from_prev(r.sup * w[s]);
      if (s == 0) {
        const double lw = r.leaf * w[0];
        const double l_dg = readlane(lw, L_DG), l_tg = readlane(lw, L_TG);
        if (lane == L_TG) o[0] += l_dg;
        if (lane == L_DG) o[0] += l_tg;
      }
      cc[0][s] = r.ch;
      cc[1][s] = r.ceta;
      cc[2][s] = r.cpi;
      cc[3][s] = r.c1;
    }
  }
  wave_sum<7>(part);
  RowSums rs{part[0], part[1], part[2], part[3], part[4], part[5], part[6]};
  if constexpr (TAPPLY) {
    const double wahp = readlane(w[0], L_AHP);
#pragma unroll
    for (int s = 0; s < NS; ++s) out[s] = o[s] + cc[0][s] * rs.sh + cc[1][s] * rs.se + cc[2][s] * rs.sp + cc[3][s] * wahp;
  }
  return rs;
}
template <int NS, bool TAPPLY, bool DFA = false>
DEB_OUTLINE RowSums coef_pass(Ctx &cx, const uint32_t (&rd)[NS], double a, double invtau,
                              const double (&w_in)[NS], const double (&Y_in)[NS], double (&out)[NS],
                              double (&fv_out)[FT_N], double *dfa_out = nullptr) {
  return coef_pass_body<NS, TAPPLY, DFA>(cx, rd, a, invtau, w_in, Y_in, out, fv_out, dfa_out);
}

// Partials of the rows' sum_fields mul[q][f] coef_f with respect to the
// features (accumulated into fb[q]) and to their bin velocity (vb[q]), for K
// contractions at once; fv[q]: the feature values of contraction q.  The
// partials of row_visit's formulas, organised for the wave rather than as one
// divergent switch over row types:
//   chain_feature_adj    the rows of the hierarchies (any slot, any lane):
//                        per-lane arithmetic with selects, no branches;
//   special_feature_adj  the rows that sit on fixed lanes of slot 0 (eta,
//                        delta_c, theta_c, delta_b, theta_b, delta_gamma,
//                        theta_gamma, delta_Q, theta_Q): their multipliers are
//                        read from those lanes and their partials formed as
//                        wave-uniform values, to be added once (on lane 0).
template <int K, class T = double>
__device__ __forceinline__ void chain_feature_adj(uint32_t d, const Ctx &cx, const T (&fv)[K][FT_N],
                                                  const T (&vq)[K], const T (&mul)[K][11], T (&fb)[K][FT_N],
                                                  T (&vb)[K]) {
  const double k = cx.k, k2 = k * k;
  const double hk2 = 0.5 / k2;
  const uint32_t t = rd_type(d);
  const double l = (double)rd_l(d);
  // opacity through the diagonal (-opac) and the polarization source (c opac)
  const bool od = t == R_F2 || t == R_FL || t == R_FLMAX || t == R_G0 || t == R_GL || t == R_G2 || t == R_GLMAX;
  const double cu = (t == R_F2 || t == R_G2) ? 0.1 : (t == R_G0 ? 0.5 : 0.0);
  // massive-nu rows: sub, sup proportional to k v_q
  const bool qr = t >= R_Q0 && t <= R_QLMAX;
  const double il = rcp_nr(2.0 * l + 1.0);
  const double cs = t == R_Q1 ? 1.0 / 3.0 : (t == R_Q2 ? 0.4 : (t == R_QL ? l * il : (t == R_QLMAX ? 1.0 : 0.0)));
  const double cp = t == R_Q0 ? 1.0 : (t == R_Q1 ? 2.0 / 3.0 : (t == R_Q2 ? 0.6 : (t == R_QL ? (l + 1.0) * il : 0.0)));
  const bool q0 = t == R_Q0, q1 = t == R_Q1, dr = t == R_DR, tr = t == R_TR;
  const double wq = (q0 || q1) ? cx.w[rd_q(d)] : 0.0;
  const T zero = tconst<T>(0.0);
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const T *m = mul[q];
    const T *F = fv[q];
    if (od) fb[q][FT_OPAC] += -m[C_DIAG] + (cu != 0.0 ? cu * m[C_UPI] : zero);
    if (qr) vb[q] += k * (m[C_SUB] * cs - m[C_SUP] * cp);
    if (q0) {
      // ch = RN wq / v * iaH; c1 = -a (RN wq / v + RN wq v)
      const T RN = F[FT_RN];
      const T iv = rcp_t<T>(vq[q]);
      fb[q][FT_RN] += m[C_CH] * wq * iv * F[FT_IAH] - m[C_C1] * F[FT_A] * wq * (iv + vq[q]);
      fb[q][FT_IAH] += m[C_CH] * RN * wq * iv;
      vb[q] += -m[C_CH] * RN * wq * F[FT_IAH] * iv * iv - m[C_C1] * F[FT_A] * RN * wq * (1.0 - iv * iv);
      fb[q][FT_A] += -m[C_C1] * RN * wq * (iv + vq[q]);
    }
    if (q1) fb[q][FT_RN] += m[C_CETA] * k * wq * hk2;
    if (dr) {
      fb[q][FT_RR] += m[C_CH] * F[FT_IAH] - 2.0 * m[C_C1] * F[FT_A];
      fb[q][FT_IAH] += m[C_CH] * F[FT_RR];
      fb[q][FT_A] += -2.0 * m[C_C1] * F[FT_RR];
    }
    if (tr) fb[q][FT_RR] += m[C_CETA] * (4.0 / 3.0) * hk2;
  }
}

// the fixed-lane rows of slot 0 (multipliers ms0 of slot 0): wave-uniform
// partials added into sp
template <int K, class T = double>
__device__ __forceinline__ void special_feature_adj(const Ctx &cx, const T (&fv)[K][FT_N], const T (&ms0)[K][11],
                                                    T (&sp)[K][FT_N]) {
  const double k = cx.k, k2 = k * k;
  const double ik2 = 1.0 / k2, hk2 = 0.5 * ik2;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const T *F = fv[q];
    auto M = [&](int lane, int f) { return t_readlane(ms0[q][f], lane); };
    T *o = sp[q];
    // eta: ch = 2 k^2 iaH
    o[FT_IAH] += M(L_ETA, C_CH) * 2.0 * k2;
    // delta_c, delta_b: ch = R iaH, c1 = -a R (R = rho_c / a, rho_b / a)
    {
      const T ch = M(L_DC, C_CH), c1 = M(L_DC, C_C1);
      o[FT_RC] += ch * F[FT_IAH] - c1 * F[FT_A];
      o[FT_IAH] += ch * F[FT_RC];
      o[FT_A] += -c1 * F[FT_RC];
    }
    {
      const T ch = M(L_DB, C_CH), c1 = M(L_DB, C_C1);
      o[FT_RB] += ch * F[FT_IAH] - c1 * F[FT_A];
      o[FT_IAH] += ch * F[FT_RB];
      o[FT_A] += -c1 * F[FT_RB];
    }
    // theta_c: diag = -aH, ceta = R_c hk2
    o[FT_AH] += -M(L_TC, C_DIAG);
    o[FT_RC] += M(L_TC, C_CETA) * hk2;
    // theta_b: sub = k^2 cs2, diag = -aH - pb43 opac, sup = pb43 opac, ceta = R_b hk2
    {
      const T sub = M(L_TB, C_SUB), dg = M(L_TB, C_DIAG), sup = M(L_TB, C_SUP);
      o[FT_CS2] += sub * k2;
      o[FT_AH] += -dg;
      o[FT_PB43] += (sup - dg) * F[FT_OPAC];
      o[FT_OPAC] += (sup - dg) * F[FT_PB43];
      o[FT_RB] += M(L_TB, C_CETA) * hk2;
    }
    // delta_gamma: ch = R_g iaH, c1 = -2 a R_g
    {
      const T ch = M(L_DG, C_CH), c1 = M(L_DG, C_C1);
      o[FT_RG] += ch * F[FT_IAH] - 2.0 * c1 * F[FT_A];
      o[FT_IAH] += ch * F[FT_RG];
      o[FT_A] += -2.0 * c1 * F[FT_RG];
    }
    // theta_gamma: sub = opac, diag = -opac, ceta = (4/3) R_g hk2
    o[FT_OPAC] += M(L_TG, C_SUB) - M(L_TG, C_DIAG);
    o[FT_RG] += M(L_TG, C_CETA) * (4.0 / 3.0) * hk2;
    if (!cx.lambda) {
      const T w = F[FT_W], ca2 = F[FT_CA2], cs2Q = F[FT_CS2Q], RQ = F[FT_RQ];
      const T opw = 1.0 + w;
      {  // delta_Q
        const T dg = M(L_DQ, C_DIAG), sup = M(L_DQ, C_SUP), uh = M(L_DQ, C_UH), ch = M(L_DQ, C_CH),
                c1 = M(L_DQ, C_C1);
        const T g = F[FT_AH] * F[FT_AH] * ik2;
        // diag = -3 (cs2Q - w) aH
        o[FT_CS2Q] += -3.0 * F[FT_AH] * dg;
        o[FT_W] += 3.0 * F[FT_AH] * dg;
        o[FT_AH] += -3.0 * (cs2Q - w) * dg;
        // sup = -opw - 9 opw (cs2Q - ca2) aH^2 / k2
        o[FT_W] += sup * (-1.0 - 9.0 * (cs2Q - ca2) * g);
        o[FT_CS2Q] += -9.0 * opw * g * sup;
        o[FT_CA2] += 9.0 * opw * g * sup;
        o[FT_AH] += -18.0 * opw * (cs2Q - ca2) * F[FT_AH] * ik2 * sup;
        // uh = -0.5 opw
        o[FT_W] += -0.5 * uh;
        // ch = RQ iaH
        o[FT_RQ] += ch * F[FT_IAH];
        o[FT_IAH] += ch * RQ;
        // c1 = -a (RQ + 3 cs2Q RQ)
        o[FT_A] += -c1 * RQ * (1.0 + 3.0 * cs2Q);
        o[FT_RQ] += -c1 * F[FT_A] * (1.0 + 3.0 * cs2Q);
        o[FT_CS2Q] += -c1 * 3.0 * F[FT_A] * RQ;
      }
      {  // theta_Q
        const T sub = M(L_TQ, C_SUB), dg = M(L_TQ, C_DIAG), ce = M(L_TQ, C_CETA), c1 = M(L_TQ, C_C1);
        // sub = cs2Q / opw k2
        const T iopw = rcp_t<T>(opw);
        o[FT_CS2Q] += sub * k2 * iopw;
        o[FT_W] += -sub * cs2Q * k2 * (iopw * iopw);
        // diag = -(1 - 3 cs2Q) aH
        o[FT_CS2Q] += 3.0 * F[FT_AH] * dg;
        o[FT_AH] += -(1.0 - 3.0 * cs2Q) * dg;
        // ceta = opw RQ 0.5 / k2
        o[FT_W] += ce * RQ * hk2;
        o[FT_RQ] += ce * opw * hk2;
        // c1 = -a 9 (cs2Q - ca2) aH opw RQ / k2
        const double c9 = 9.0 * ik2;
        o[FT_A] += -c1 * c9 * (cs2Q - ca2) * F[FT_AH] * opw * RQ;
        o[FT_CS2Q] += -c1 * F[FT_A] * c9 * F[FT_AH] * opw * RQ;
        o[FT_CA2] += c1 * F[FT_A] * c9 * F[FT_AH] * opw * RQ;
        o[FT_AH] += -c1 * F[FT_A] * c9 * (cs2Q - ca2) * opw * RQ;
        o[FT_W] += -c1 * F[FT_A] * c9 * (cs2Q - ca2) * F[FT_AH] * RQ;
        o[FT_RQ] += -c1 * F[FT_A] * c9 * (cs2Q - ca2) * F[FT_AH] * opw;
      }
    }
  }
}

// First-order contraction of sum_rows w_row f_row(a, Y) along (a, theta):
//   slot_muls          the multiplier of every coefficient field of a row;
//   feature_partials   per lane, the partials of its rows' sum_fields
//                      mul * coef with respect to the features, accumulated
//                      into fb (plain doubles, no cross-lane sums; linear in
//                      mul, so contractions at the same features may share fb);
//   lane_dot           after fb is summed over the wave, sum_m dF_m fb_m for
//                      the lane's own tangent direction (lane_ctx).
// The functionals (h', eta', Pi, f_1) enter through their double values in rs,
// which is exact at first order: d(u H) = du H + u dH with dH = sum_rows dc_h Y,
// so sum_rows w u dH = s_h sum_rows dc_h Y (multiplier Y s_h of the c_h field).

// multipliers of the coefficient fields of slot s's rows in
// sum_rows w f_row(Y) (the contraction's per-row weights): ys, ws the slot's
// values, ym, yp its neighbours' (lanes i-1, i+1), ydg, ytg, wahp the
// wave-uniform Y[L_DG], Y[L_TG], w[L_AHP] of slot 0; rs the functionals' values
template <class T>
__device__ __forceinline__ void slot_muls(int s, double ys, double ws, double ym, double yp, double ydg, double ytg,
                                          double wahp, const RowSumsT<T> &rs, T (&ms)[11]) {
  const int lane = lane_id();
  const double lp = (s == 0) ? (lane == L_DG ? ytg : (lane == L_TG ? ydg : 0.0)) : 0.0;
  ms[C_SUB] = tconst<T>(ws * ym);
  ms[C_DIAG] = tconst<T>(ws * ys);
  ms[C_SUP] = tconst<T>(ws * yp);
  ms[C_UH] = ws * rs.H;
  ms[C_UETA] = ws * rs.E;
  ms[C_UPI] = ws * rs.Pi;
  ms[C_CH] = ys * rs.sh;
  ms[C_CETA] = ys * rs.se;
  ms[C_CPI] = ys * rs.sp;
  ms[C_C1] = tconst<T>(ys * wahp);
  // the leaf field is unused by the a-row: there it carries the row weight w_a
  ms[C_LEAF] = tconst<T>((s == 0 && lane == L_A) ? ws : ws * lp);
}

// fb[q] += partials, with respect to the features fv[q], of
// sum_rows sum_fields mul_q coef_field, mul_q the row multipliers of
// sum_rows w[q] f_row(Y[q]) (slot_muls, formed slot by slot here; rs[q] the
// functionals' values): K contractions in one pass of the divergent row switch
template <int NS, int K, class T = double>
__device__ __forceinline__ void feature_partials_body(const Ctx &cx, const uint32_t (&rd)[NS],
                                                      const T (&fv_in)[K][FT_N], const double (&Y_in)[K][NS],
                                                      const double (&w_in)[K][NS], const RowSumsT<T> (&rs_in)[K],
                                                      T (&fb)[K][FT_N]) {
  const int lane = lane_id();
  // Private copies: fb may alias the inputs as far as the compiler knows, so
  // accumulating through the references would reload every operand after each
  // store (memory traffic in the caller's frame); locals stay in registers.
  T fv[K][FT_N], acc[K][FT_N];
  double Y[K][NS], w[K][NS];
  RowSumsT<T> rs[K];
  double ydg[K], ytg[K], wahp[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
#pragma unroll
    for (int m = 0; m < FT_N; ++m) {
      fv[q][m] = fv_in[q][m];
      acc[q][m] = tconst<T>(0.0);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      Y[q][s] = Y_in[q][s];
      w[q][s] = w_in[q][s];
    }
    rs[q] = rs_in[q];
    ydg[q] = readlane(Y[q][0], L_DG);
    ytg[q] = readlane(Y[q][0], L_TG);
    wahp[q] = readlane(w[q][0], L_AHP);
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t ty = rd_type(rd[s]);
    const bool qrow = ty >= R_Q0 && ty <= R_QLMAX;
    const double iq = qrow ? rcp_nr(cx.q[rd_q(rd[s])]) : 0.0;
    T ms[K][11], vq[K], xq[K], vb[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      slot_muls<T>(s, Y[q][s], w[q][s], from_prev(Y[q][s]), from_next(Y[q][s]), ydg[q], ytg[q], wahp[q], rs[q],
                   ms[q]);
      // v_q(x), x = a amnu / q
      xq[q] = fv[q][FT_A] * fv[q][FT_AMNU] * iq;
      vq[q] = rsqrt_t<T>(1.0 + xq[q] * xq[q]);
      vb[q] = tconst<T>(0.0);
    }
    chain_feature_adj<K, T>(rd[s], cx, fv, vq, ms, acc, vb);
    if (s == 0) {
      T sp[K][FT_N];
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int m = 0; m < FT_N; ++m) sp[q][m] = tconst<T>(0.0);
      special_feature_adj<K, T>(cx, fv, ms, sp);
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q)
#pragma unroll
          for (int m = 0; m < FT_N; ++m) acc[q][m] += sp[q][m];
      }
    }
    if (qrow) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T dvdx = -xq[q] * vq[q] * vq[q] * vq[q];
        acc[q][FT_A] += vb[q] * dvdx * fv[q][FT_AMNU] * iq;
        acc[q][FT_AMNU] += vb[q] * dvdx * fv[q][FT_A] * iq;
      }
    }
  }
  if (lane == L_A) {  // f_a = aH a, weight w_a (the leaf multiplier of the a-row)
#pragma unroll
    for (int q = 0; q < K; ++q) {
      acc[q][FT_AH] += w[q][0] * fv[q][FT_A];
      acc[q][FT_A] += w[q][0] * fv[q][FT_AH];
    }
  }
#pragma unroll
  for (int q = 0; q < K; ++q)
#pragma unroll
    for (int m = 0; m < FT_N; ++m) fb[q][m] += acc[q][m];
}
template <int NS, int K, class T = double>
DEB_OUTLINE void feature_partials(const Ctx &cx, const uint32_t (&rd)[NS], const T (&fv_in)[K][FT_N],
                                  const double (&Y_in)[K][NS], const double (&w_in)[K][NS],
                                  const RowSumsT<T> (&rs_in)[K], T (&fb)[K][FT_N]) {
  feature_partials_body<NS, K, T>(cx, rd, fv_in, Y_in, w_in, rs_in, fb);
}

// The feature partials of one reverse stage at five slots (after its row pass
// and row sums): both contractions summed over the wave, the stage's (at its
// features fv) stored to fb_out and contracted with the features'
// a-derivatives dfa into the returned a-direction of the stage cotangent, the
// J-term's (at the features fvJ of a0) accumulated into fbJ.  fvJ, fbJ and
// fb_out: one copy per wave in LDS.
template <int NS>
DEB_OUTLINE double stage_partials(const Ctx &cx, const uint32_t (&rd)[NS], const double (&fv_in)[FT_N],
                                  const lds_double *fvJ_in, const double (&dfa_in)[FT_N],
                                  const double (&Y_in)[2][NS], const double (&w_in)[2][NS],
                                  const RowSums (&rs_in)[2], lds_double *fbJ_io, lds_double *fb_out) {
  double fvs[2][FT_N], fbs[2][FT_N];
#pragma unroll
  for (int q = 0; q < FT_N; ++q) {
    fvs[0][q] = fv_in[q];
    fvs[1][q] = (double)fvJ_in[q];
    fbs[0][q] = 0.0;
    fbs[1][q] = 0.0;
  }
  feature_partials_body<NS, 2>(cx, rd, fvs, Y_in, w_in, rs_in, fbs);
  wave_sum<FT_N>(fbs[0]);
  wave_sum<FT_N>(fbs[1]);
  double ga = 0.0;
#pragma unroll
  for (int q = 0; q < FT_N; ++q) ga += dfa_in[q] * fbs[0][q];
#pragma unroll
  for (int q = 0; q < FT_N; ++q) {
    fb_out[q] = fbs[0][q];
    fbJ_io[q] += fbs[1][q];
  }
  return ga;
}

// One tangent direction per lane: lane & 7 == 0 carries d/da (component 0 of
// TL, seeded by seed_tl), lane & 7 == c (1 <= c <= ndir) the input tangent
// c - 1 (component 1, loaded through lane_ctx); TL arithmetic then costs
// three doubles per value instead of 2 + MAXDIR.  Only wave-uniform
// quantities may be computed in TL (a lane sum would mix directions).
using TL = dn<2>;

__device__ __forceinline__ Ctx lane_ctx(const Ctx &cx) {
  Ctx cl = cx;
  const int c = lane_id() & 7;
  const bool on = c >= 1 && c <= cx.ndir;
  cl.ndir = on ? 1 : 0;
  cl.dcos = cx.dcos + (on ? (c - 1) * COSMO_DOUBLES : 0);
  cl.dtab = cx.dtab + (on ? (c - 1) * cx.stride : 0);
  return cl;
}

__device__ __forceinline__ TL seed_tl(double a) {
  TL r;
  r = a;
  r.d[0] = (lane_id() & 7) == 0 ? 1.0 : 0.0;
  return r;
}

// the lane's direction of sum_m F_m fbs_m (fbs wave-uniform)
__device__ __forceinline__ double lane_dot(const TL (&F)[FT_N], const double (&fbs)[FT_N]) {
  double g = 0.0;
#pragma unroll
  for (int m = 0; m < FT_N; ++m) g += (F[m].d[0] + F[m].d[1]) * fbs[m];
  return g;
}

// feature values at a (plain doubles), wave-uniform: stored once per wave in
// LDS (every lane writes the same values)
DEB_OUTLINE void feature_values(Ctx &cx, double a, lds_double *fv) {
  double f[FT_N];
  features_t<double>(cx, a, f);
#pragma unroll
  for (int m = 0; m < FT_N; ++m) fv[m] = f[m];
}

// the lane's direction of sum_m F_m(a) fbs_m (fbs wave-uniform): d/da on
// lanes with lane & 7 == 0, d/dtheta_{c-1} on lanes with lane & 7 == c
DEB_OUTLINE double lane_contract(const Ctx &cx, double a, const double (&fbs)[FT_N]) {
  Ctx cl = lane_ctx(cx);
  TL F[FT_N];
  features_t<TL>(cl, seed_tl(a), F);
  return lane_dot(F, fbs);
}

// The input-tangent directions of sum_m F_m(a_i) fbs_im for the eight stages
// of a reverse step at once: lane l evaluates the features at stage l >> 3's
// a along direction (l & 7) - 1 (none on l & 7 == 0).  cst: the stages' a_i
// (8 doubles), then their wave-uniform partials fbs_i (8 x FT_N).
DEB_OUTLINE double lane_contract_stages(const Ctx &cx, const lds_double *cst) {
  Ctx cl = lane_ctx(cx);
  const int i = lane_id() >> 3;
  TL as;
  as = (double)cst[i];
  TL F[FT_N];
  features_t<TL>(cl, as, F);
  double g = 0.0;
#pragma unroll
  for (int m = 0; m < FT_N; ++m) g += F[m].d[1] * (double)cst[8 + i * FT_N + m];
  return g;
}

// The a-column pass of the reverse step in dual arithmetic along a, with one
// background evaluation for all of it: out = A(a)^T w for the y-linear part of
// f (rows other than a; the a-column and the a-row go through the features)
// and its a-derivative; then, for S = sum_rows w f_row(a, Y), the functional
// sums and the per-lane feature partials P (fb.v) with their a-derivatives Q
// (fb.d), at the features of a.
template <int NS>
DEB_OUTLINE void acol_pass(Ctx &cx, const uint32_t (&rd)[NS], dual a, double invtau, const double (&w_in)[NS],
                           const double (&Y_in)[NS], dual (&out_io)[NS], dual (&fb_out)[FT_N]) {
  using T = dual;
  const int lane = lane_id();
  double w[NS], Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    w[s] = w_in[s];
    Y[s] = Y_in[s];
  }
  const Bg<T> b = background<T>(cx, a);
  T F[1][FT_N];
  features_of<T>(cx, a, b, F[0]);
  T out[NS], part[7], cc[4][NS];
#pragma unroll
  for (int i = 0; i < 7; ++i) part[i] = tconst<T>(0.0);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const Coef<T> r = row_coef_slot<T>(rd[s], b, cx, invtau, s == 0);
    T sw = r.sub * w[s];
    T pw = r.sup * w[s];
    out[s] = r.diag * w[s] + t_from_next(sw) + t_from_prev(pw);
    if (s == 0) {
      T lw = r.leaf * w[0];
      T l_dg = t_readlane(lw, L_DG), l_tg = t_readlane(lw, L_TG);
      if (lane == L_TG) out[0] = out[0] + l_dg;
      if (lane == L_DG) out[0] = out[0] + l_tg;
    }
    part[0] = part[0] + r.ch * Y[s];
    part[1] = part[1] + r.ceta * Y[s];
    part[2] = part[2] + r.cpi * Y[s];
    part[3] = part[3] + r.c1 * Y[s];
    part[4] = part[4] + r.uh * w[s];
    part[5] = part[5] + r.ueta * w[s];
    part[6] = part[6] + r.upi * w[s];
    cc[0][s] = r.ch;
    cc[1][s] = r.ceta;
    cc[2][s] = r.cpi;
    cc[3][s] = r.c1;
  }
  t_wave_sum<T, 7>(part);
  const double wahp = readlane(w[0], L_AHP);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    out_io[s] = out[s] + cc[0][s] * part[4] + cc[1][s] * part[5] + cc[2][s] * part[6] + cc[3][s] * wahp;
  const RowSumsT<T> rs[1] = {RowSumsT<T>{part[0], part[1], part[2], part[3], part[4], part[5], part[6]}};
  double Yd[1][NS], wd[1][NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Yd[0][s] = Y[s];
    wd[0][s] = w[s];
  }
  T fb[1][FT_N];
#pragma unroll
  for (int q = 0; q < FT_N; ++q) fb[0][q] = tconst<T>(0.0);
  feature_partials_body<NS, 1, T>(cx, rd, F, Yd, wd, rs, fb);
#pragma unroll
  for (int q = 0; q < FT_N; ++q) fb_out[q] = fb[0][q];
}

// Second order along one direction per lane: with P_m = dS/dF_m and
// Q_m = dP_m/da (wave-uniform) for S(F(a, theta)), returns on lane direction c
// d/dc (dS/da) = sum_m Q_m dF_m/dc + P_m d2F_m/(da dc)  (c = a on lane & 7 == 0).
// PQ: P (FT_N doubles), then Q, one copy per wave in LDS.  Features in nested
// duals: outer component along the lane's direction, inner along a.
using TL2 = dn<2, dual>;
DEB_OUTLINE double lane_contract2(const Ctx &cx, double a, const lds_double *PQ) {
  Ctx cl = lane_ctx(cx);
  TL2 as;
  as = a;
  as.v.d = 1.0;
  as.d[0].v = (lane_id() & 7) == 0 ? 1.0 : 0.0;
  TL2 F[FT_N];
  features_t<TL2>(cl, as, F);
  double g = 0.0;
#pragma unroll
  for (int m = 0; m < FT_N; ++m)
    g += (double)PQ[FT_N + m] * (F[m].d[0].v + F[m].d[1].v) + (double)PQ[m] * (F[m].d[0].d + F[m].d[1].d);
  return g;
}

// the lane-direction result of every direction: out[j] = d/dtheta_j
__device__ __forceinline__ void lane_gather(double g, int ndir, double (&out)[MAXDIR]) {
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j) out[j] = j < ndir ? readlane(g, j + 1) : 0.0;
}

// row sums of the step Jacobian J(t0, y0) kept by the factorisation
template <int NS>
__device__ __forceinline__ RowSums jac_sums(const Mode<NS> &m, const double (&Y)[NS], const double (&w)[NS]) {
  double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    part[0] += m.Cf[0][s] * Y[s];
    part[1] += m.Cf[1][s] * Y[s];
    part[2] += m.Cf[2][s] * Y[s];
    part[3] += m.c1[s] * Y[s];
    part[4] += m.U[0][s] * w[s];
    part[5] += m.U[1][s] * w[s];
    part[6] += m.U[2][s] * w[s];
  }
  wave_sum<7>(part);
  return RowSums{part[0], part[1], part[2], part[3], part[4], part[5], part[6]};
}

// The row work of one reverse stage as a single call: the transposed row pass
// at the stage point (ub = J(t_i, u_i)^T rb, the stage's features and their
// a-derivative dfa), the row sums of the J(t0, y0) k_i term, and the feature
// partials of both contractions, each summed over the wave: the J-term's (at
// the features fvJ of a0, multipliers without the a-row) accumulated into fbJ,
// the stage's stored to fb_out and contracted with dfa into the a-direction of
// ub (fbJ and fb_out: one copy per wave in LDS).
template <int NS>
DEB_OUTLINE void stage_rows(Ctx &cx, const Mode<NS> &m, double ai, double iti, const double (&rb_in)[NS],
                            const double (&u_in)[NS], const double (&ki_in)[NS], const lds_double *fvJ_in,
                            double (&ub_out)[NS], lds_double *fbJ_io, lds_double *fb_out) {
  const int lane = lane_id();
  double rb[NS], u[NS], ki[NS], ub[NS], dfa[FT_N];
  double fvs[2][FT_N], fbs[2][FT_N];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rb[s] = rb_in[s];
    u[s] = u_in[s];
    ki[s] = ki_in[s];
  }
#pragma unroll
  for (int q = 0; q < FT_N; ++q) {
    fvs[1][q] = (double)fvJ_in[q];
    fbs[0][q] = 0.0;
    fbs[1][q] = 0.0;
  }
  RowSums rsq[2];
  rsq[0] = coef_pass_body<NS, true, true>(cx, m.rd, ai, iti, rb, u, ub, fvs[0], dfa);
  double Yq[2][NS], wq[2][NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Yq[0][s] = u[s];
    wq[0][s] = rb[s];
    Yq[1][s] = ki[s];
    wq[1][s] = (s == 0 && lane == L_A) ? 0.0 : rb[s];
  }
  rsq[1] = jac_sums<NS>(m, ki, wq[1]);
  feature_partials_body<NS, 2>(cx, m.rd, fvs, Yq, wq, rsq, fbs);
  wave_sum<FT_N>(fbs[0]);
  wave_sum<FT_N>(fbs[1]);
  double ga = 0.0;
#pragma unroll
  for (int q = 0; q < FT_N; ++q) ga += dfa[q] * fbs[0][q];
  if (lane == L_A) ub[0] += ga;
#pragma unroll
  for (int s = 0; s < NS; ++s) ub_out[s] = ub[s];
#pragma unroll
  for (int q = 0; q < FT_N; ++q) {
    fb_out[q] = fbs[0][q];
    fbJ_io[q] += fbs[1][q];
  }
}

// per-lane multiplier of the explicit tau dependence (hierarchy truncations):
// d f_row / d tau = trunc_row / tau^2 * y_row with trunc_row = l + 1
template <int NS>
__device__ __forceinline__ void trunc_weights(const uint32_t (&rd)[NS], double (&tw)[NS]) {
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t t = rd_type(rd[s]);
    const bool trunc = (t == R_FLMAX) | (t == R_GLMAX) | (t == R_FRLMAX) | (t == R_QLMAX);
    tw[s] = trunc ? (rd_l(rd[s]) + 1.0) : 0.0;
  }
}

__device__ __forceinline__ TD seed_td(double v, double d0) {
  TD r;
  r = v;
  r.d[0] = d0;
  return r;
}

// one forward step from (t0, y0) to y1, its eight stage vectors stored in
// K[8][NS][64] (inlined, with its factorisation, right-hand sides and solves
// as calls: the whole step as one call is large enough to need long branches,
// which must stay out of non-kernel functions)
template <int NS>
__device__ __forceinline__ void step_store(Mode<NS> &m, Ctx &cx, double t0, double h, const double (&y0)[NS],
                                           double (&y1)[NS], double *K, lds_double *kl) {
  const int lane = lane_id();
  double err[NS];
  auto store = [&](int i, const double (&k)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) K[(i * NS + s) * WAVE + lane] = k[s];
  };
  rodas5_step_s<NS, decltype(store), true>(m, cx, t0, h, y0, y1, err, store, kl);
}

// ---------------------------------------------------------------------------
// reverse of one Rodas5 step.  In: ybar = dL/dy1.  Out: ybar = dL/dy0,
// tbar += dL/dt0 (h fixed), hbar += dL/dh, thb[j] += dL/dtheta_j (wave-uniform).
// K: the step's eight stage vectors ([8][NS][64], written by step_store);
// KBl: per-wave LDS scratch of 7 stage vectors.  The factorisation of W at
// (t0, y0) is recomputed into m.
template <int NS>
__device__ __forceinline__ void rodas5_step_vjp(Mode<NS> &m, Ctx &cx, double t0, double h,
                                                const double (&y0)[NS], double (&ybar)[NS], double (&thb)[MAXDIR],
                                                double &tbar, double &hbar, const double *K, double *KBl) {
  const int lane = lane_id();
  auto Kp = [&](int i, int s) -> double { return K[(i * NS + s) * WAVE + lane]; };
  auto KBp = [&](int i, int s) -> double & { return KBl[(i * NS + s) * WAVE + lane]; };
  // per stage: its a and its wave-uniform feature partials (written by every
  // lane alike)
  __shared__ double cst_[8 * (1 + FT_N)];
  lds_double *const cst = (lds_double *)cst_;
  factor_call<NS>(m, cx, t0, y0, h);
  double tw[NS];
  trunc_weights<NS>(m.rd, tw);
  const double it0 = 1.0 / t0;
  const double a0 = readlane(y0[0], L_A);
  double dT[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) dT[s] = tw[s] * it0 * it0 * y0[s];
  const double gamma = R5::gamma;
  const double ih = 1.0 / h;
  double y1b[NS], y0b[NS], dTb[NS], rho[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    y1b[s] = ybar[s];
    y0b[s] = ybar[s];  // y1 = y0 + sum A8 k
    dTb[s] = 0.0;
    rho[s] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j)
#pragma unroll
    for (int s = 0; s < NS; ++s) KBp(j, s) = kA8[j] * y1b[s];
  double th_acc[MAXDIR];
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j) th_acc[j] = 0.0;
  double a0b = 0.0, t0b = 0.0, hb = 0.0;
  // Feature partials, one row pass per stage (stage_rows) with two
  // contractions: f(t_i, u_i) at the stage's features (contracted every
  // stage) and the J(t0, y0) k_i term at the features fvJ of a0 (its partials
  // fbJ accumulate over the stages and are contracted once).
  // the features of a0 and the J-term's partials (summed over the wave) are
  // wave-uniform: one copy per wave in LDS (written by every lane alike)
  __shared__ double fvJ_[FT_N], fbJ_[FT_N];
  lds_double *const fvJ = (lds_double *)fvJ_, *const fbJl = (lds_double *)fbJ_;
  feature_values(cx, a0, fvJ);
#pragma unroll
  for (int q = 0; q < FT_N; ++q) fbJl[q] = 0.0;
#pragma nounroll
  for (int i = 7; i >= 0; --i) {
    double rb[NS], ki[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) rb[s] = (i == 7) ? y1b[s] * kA8[7] : KBp(i, s);
    wsolveT<NS>(m, rb);
    // stage point u_i and the combination sum_j C_ij k_j
    double u[NS], ck[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      u[s] = y0[s];
      ck[s] = 0.0;
    }
    // this stage's vector and the earlier ones in one batch of loads (the
    // tableau's rows are zero from column i on, so the sums over all seven add
    // exact zeros)
    {
      double kj[7][NS];
#pragma unroll
      for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int s = 0; s < NS; ++s) kj[j][s] = Kp(j, s);
#pragma unroll
      for (int s = 0; s < NS; ++s) ki[s] = Kp(i, s);
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const double aij = kA[i][j], cij = kC[i][j];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          u[s] += aij * kj[j][s];
          ck[s] += cij * kj[j][s];
        }
      }
    }
    for (int j = 0; j < i; ++j) {
      const double cij = kC[i][j];
#pragma unroll
      for (int s = 0; s < NS; ++s) KBp(j, s) += (cij * ih) * rb[s];
    }
    const double ti = t0 + kc[i] * h;
    const double iti = 1.0 / ti;
    // dot products: h-derivative of W, C/h and d h dT terms; explicit time
    // dependence at the stage point and of J at (t0, y0)
    double dots[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      dots[0] += rb[s] * (ki[s] / (gamma * h * h) - ck[s] * ih * ih + kd[i] * dT[s]);
      dots[1] += rb[s] * tw[s] * iti * iti * u[s];
      dots[2] += rb[s] * tw[s] * it0 * it0 * ki[s];
    }
    wave_sum<3>(dots);
    hb += dots[0] + kc[i] * dots[1];
    t0b += dots[1] + dots[2];
    if (kd[i] != 0.0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) dTb[s] += kd[i] * h * rb[s];
    }
    // f(t_i, u_i): ubar = J(t_i, u_i)^T rbar, and the J(t0, y0) k_i term
    // (its y-linear part at a0; the a-column goes through rho)
    double ub[NS];
    {
      const double ai = readlane(u[0], L_A);
      if constexpr (NS <= 4) {
        stage_rows<NS>(cx, m, ai, iti, rb, u, ki, fvJ, ub, fbJl, cst + 8 + i * FT_N);
      } else {
        // five slots: the same work as separate calls (as one call the
        // reverse launch of the lmax-50 workload measured 14 % slower)
        double fvs[FT_N], dfl[FT_N], Yq[2][NS], wq[2][NS];
        RowSums rsq[2];
        rsq[0] = coef_pass<NS, true, true>(cx, m.rd, ai, iti, rb, u, ub, fvs, dfl);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          Yq[0][s] = u[s];
          wq[0][s] = rb[s];
          Yq[1][s] = ki[s];
          wq[1][s] = (s == 0 && lane == L_A) ? 0.0 : rb[s];
        }
        rsq[1] = jac_sums<NS>(m, ki, wq[1]);
        // the a-direction now (it feeds the earlier stages); the input-tangent
        // directions of all stages in one pass after the loop
        const double ga = stage_partials<NS>(cx, m.rd, fvs, fvJ, dfl, Yq, wq, rsq, fbJl, cst + 8 + i * FT_N);
        if (lane == L_A) ub[0] += ga;
      }
      cst[i] = ai;
      const double kia = readlane(ki[0], L_A);
#pragma unroll
      for (int s = 0; s < NS; ++s) rho[s] += kia * rb[s];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) y0b[s] += ub[s];
    for (int j = 0; j < i; ++j) {
      const double aij = kA[i][j];
#pragma unroll
      for (int s = 0; s < NS; ++s) KBp(j, s) += aij * ub[s];
    }
  }
  // the stages' input-tangent directions
  {
    const double g = lane_contract_stages(cx, cst);
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) {
      if (j < cx.ndir) {
        double t = 0.0;
#pragma unroll
        for (int st = 0; st < 8; ++st) t += readlane(g, st * 8 + j + 1);
        th_acc[j] += t;
      }
    }
  }
  // the J(t0, y0) k_i terms' partials, contracted with the feature tangents
  // at a0 together with the a-column's below (first order: added to Q)
  double fbJ[FT_N];
#pragma unroll
  for (int q = 0; q < FT_N; ++q) fbJ[q] = fbJl[q];
  // a-column of J(t0, y0): rho^T df/da, differentiated along y0, a0 and theta
  {
    dual ad;
    ad.v = a0;
    ad.d = 1.0;
    // rho^T A'(a) (the derivative of the transposed apply), and for
    // S = rho^T f(a, y0) the per-lane feature partials in dual arithmetic
    dual gy[NS], fbd[FT_N];
    acol_pass<NS>(cx, m.rd, ad, it0, rho, y0, gy, fbd);
#pragma unroll
    for (int s = 0; s < NS; ++s) y0b[s] += gy[s].d;
    // the feature partials P and their a-derivative Q, summed over the wave
    // and kept in the stage block (whose contraction has been read), then the
    // features' second derivatives
    double PQ[2 * FT_N];
#pragma unroll
    for (int q = 0; q < FT_N; ++q) {
      PQ[q] = fbd[q].v;
      PQ[FT_N + q] = fbd[q].d;
    }
    wave_sum<2 * FT_N>(PQ);
#pragma unroll
    for (int q = 0; q < FT_N; ++q) {
      cst[q] = PQ[q];
      cst[FT_N + q] = PQ[FT_N + q] + fbJ[q];
    }
    const double g = lane_contract2(cx, a0, cst);
    a0b += readlane(g, 0);
    double tl[MAXDIR];
    lane_gather(g, cx.ndir, tl);
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) th_acc[j] += tl[j];
  }
  // dT = D(t0) y0
  {
    double pt = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      y0b[s] += tw[s] * it0 * it0 * dTb[s];
      pt += dTb[s] * (-2.0 * tw[s] * it0 * it0 * it0) * y0[s];
    }
    t0b += wave_sum1(pt);
  }
  if (lane == L_A) y0b[0] += a0b;
#pragma unroll
  for (int s = 0; s < NS; ++s) ybar[s] = y0b[s];
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j) thb[j] += th_acc[j];
  tbar += t0b;
  hbar += hb;
}

// ---------------------------------------------------------------------------
// start time and initial conditions in tangent arithmetic (the same formulas
// as start_time and initial_conditions)

template <class T>
__device__ __forceinline__ T start_time_t(Ctx &cx, T tau_out_min) {
  using P = ParT<T>;
  const P taumin = cv<T>(cx, F_TAUMIN), YHe = cv<T>(cx, F_YHE), Omegab = cv<T>(cx, F_OMEGAB), H0 = cv<T>(cx, F_H0);
  const T tau0 = taumin;
  const T tau1 = spl_at<T>(cx, DEB_SPL_TAU_A, tconst<T>(0.1));
  const double tau_k = 1.0 / cx.k;
  const P akthom = 2.3048e-9 * (1.0 - YHe) * Omegab * H0 * H0;
  T lt[2];
#pragma unroll
  for (int which = 0; which < 2; ++which) {
    // the bisection compares signs only: the root's tangent flows through
    // the bracket end points
    T xl = dlog(tau0), xr = dlog(tau1);
    auto fn = [&](const T &x) {
      T tau = dexp(x);
      T a = spl_at<T>(cx, DEB_SPL_A_TAU, tau);
      Bg<T> b = background<T>(cx, a);
      if (which == 0) return vval((1.0 / b.aH) / tau_k / 0.07 - 1.0);
      T xe = spl_at<T>(cx, DEB_SPL_XE_TAU, tau);
      T opac = xe * akthom / (a * a);
      return vval((1.0 / opac) / (1.0 / b.aH) / 0.0004 - 1.0);
    };
    for (int i = 0; i < 7; ++i) {
      T xm = 0.5 * (xl + xr);
      if (fn(xm) * fn(xl) > 0) xl = xm; else xr = xm;
    }
    lt[which] = 0.5 * (xl + xr);
  }
  T ts = dexp(vval(lt[1]) < vval(lt[0]) ? lt[1] : lt[0]);
  return 0.99 * (vval(tau_out_min) < vval(ts) ? tau_out_min : ts);
}

template <int NS, class T>
__device__ __forceinline__ void ics_t(Ctx &cx, const uint32_t (&rd)[NS], T tau, T (&y)[NS]) {
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), Neff = cv<T>(cx, F_NEFF);
  const P Nmnu = cv<T>(cx, F_NMNU), amnu = cv<T>(cx, F_AMNU), w0 = cv<T>(cx, F_W0), wa = cv<T>(cx, F_WA);
  const P cs2Q = cv<T>(cx, F_CS2DE);
  const double k = cx.k;
  T a = spl_at<T>(cx, DEB_SPL_A_TAU, tau);
  T rnu = dexp(spl_at<T>(cx, DEB_SPL_LOGRHONU_LOGA, dlog(a)));
  T a4 = a * a * a * a;
  T rhom = grhom * Omegam / (a * a * a);
  T rhor = (grhog + grhor * (Neff + Nmnu * rnu)) / a4;
  T rhonu = grhor * (Neff + Nmnu * rnu) / a4;
  P fracb = Omegab / Omegam;
  T fracnu = rhonu / rhor;
  T om = a * rhom / dsqrt(rhor);
  const double C = -1.0;
  T kt = k * tau;
  T deltag = -(kt * kt) / 3.0 * (1.0 - om * tau / 5.0) * C;
  T thetag = -(kt * kt * kt) / tau / 36.0 *
             (1.0 - 3.0 * (1.0 + 5.0 * fracb - fracnu) / 20.0 / (1.0 - fracnu) * om * tau) * C;
  T thetar = -(kt * kt * kt * kt) / tau / 36.0 / (4.0 * fracnu + 15.0) *
             (4.0 * fracnu + 11.0 + 12.0 -
              3.0 * (8.0 * fracnu * fracnu + 50.0 * fracnu + 275.0) / 20.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  T shearr = (kt * kt) / (45.0 + 12.0 * fracnu) * 2.0 *
             (1.0 + (4.0 * fracnu - 5.0) / 4.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  T wQ = w0 + wa * (1.0 - a);
  T deltaq = (kt * kt) / 4.0 * (1.0 + wQ) * (4.0 - 3.0 * cs2Q) / (4.0 - 6.0 * wQ + 3.0 * cs2Q) * C;
  T thetaq = (kt * kt * kt * kt) / tau / 4.0 * cs2Q / (4.0 - 6.0 * wQ + 3.0 * cs2Q) * C;
  if (cx.lambda) {
    deltaq = tconst<T>(0.0);
    thetaq = tconst<T>(0.0);
  }
  T eta = C * (1.0 - (kt * kt) / 12.0 / (15.0 + 4.0 * fracnu) *
                         (5.0 + 4.0 * fracnu - (16.0 * fracnu * fracnu + 280.0 * fracnu + 325.0) / 10.0 /
                                                   (2.0 * fracnu + 15.0) * tau * om));
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t t = rd_type(rd[s]);
    const int qi = (int)rd_q(rd[s]);
    T v = tconst<T>(0.0);
    switch (t) {
      case R_A: v = a; break;
      case R_ETA: v = eta; break;
      case R_DC: v = 0.75 * deltag; break;
      case R_DB: v = 0.75 * deltag; break;
      case R_TB: v = thetag; break;
      case R_DG: v = deltag; break;
      case R_TG: v = thetag; break;
      case R_DR: v = deltag; break;
      case R_TR: v = thetar; break;
      case R_FR2: v = 2.0 * shearr; break;
      case R_DQ: v = deltaq; break;
      case R_TQ: v = thetaq; break;
      case R_Q0: case R_Q1: case R_Q2: {
        T aq = a * amnu / cx.q[qi];
        T vq = 1.0 / dsqrt(1.0 + aq * aq);
        const double dl = cx.dlnf[qi];
        if (t == R_Q0) v = -0.25 * dl * deltag;
        else if (t == R_Q1) v = -dl * thetar / vq / k / 3.0;
        else v = -0.5 * dl * shearr;
        break;
      }
      default: break;
    }
    y[s] = v;
  }
}

// ---------------------------------------------------------------------------
// the gradient kernel's arguments
struct AdjArgs {
  int ndir;
  const double *dcos;     // [n_cosmo][ndir][24]
  const double *dtab;     // [n_cosmo][ndir][stride]
  const double *dtau;     // [n_cosmo][ndir][nout]
  const double *ybar;     // [n_cosmo][nk][nout][20]  dL/d(output fields)
  const int32_t *status;  // [n_cosmo][nk] of the taped forward solve
  double *grad;           // [n_cosmo][nk][ndir]
  TapeView tape;
  double *rings;          // [n_rings][ring_doubles(NS, C)] reverse-launch rings
  int32_t *ring_busy;     // [n_rings] 1 while a wave owns the ring
  const int32_t *order;   // [n_modes] launch order (mode indices, longest first), or null
  int n_rings;
  int n_in;               // state entries read by the output conversion
  int32_t in_idx[WAVE];   // their reference indices, one per lane
};

// reference indices of the state entries read by the output conversion, one
// per lane (host side: fills AdjArgs::in_idx)
inline int output_inputs(const deb_config_t &c, int32_t *idx) {
  const int lg = c.lmaxg, lgp = c.lmaxgp, lr = c.lmaxr, nq = c.nqmax;
  const int n = 12 + lg + lgp + lr + nq * (c.lmaxnu + 1);
  const int ir = 9 + lg + lgp, iq0 = 10 + lg + lgp + lr;
  int m = 0;
  for (int i : {0, 2, 3, 4, 5, 6, 7, 8}) idx[m++] = i;
  idx[m++] = ir;
  idx[m++] = ir + 1;
  for (int i = 0; i < nq; ++i) idx[m++] = iq0 + i;
  for (int i = 0; i < nq; ++i) idx[m++] = iq0 + nq + i;
  idx[m++] = n - 2;
  idx[m++] = n - 1;
  for (int i = m; i < WAVE; ++i) idx[i] = -1;
  return m;
}

__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// reverse of the output conversion of one output time: yb = dL/dy (slot
// layout), thb += dL/dtheta at fixed y.  w: the 20 field cotangents.
template <int NS>
__device__ __forceinline__ void output_vjp(const Params<NS> &P, Ctx &cx, const AdjArgs &A, const uint32_t (&rd)[NS],
                                           const double (&yo)[NS], const double *w, double *ls, double *lg,
                                           double (&yb)[NS], double (&thb)[MAXDIR]) {
  const int lane = lane_id();
  const int L0 = P.cfg.lmaxg, L1 = P.cfg.lmaxgp, L2 = P.cfg.lmaxr;
  for (int i = lane; i < P.n; i += WAVE) lg[i] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    if (rd_type(rd[s]) != R_NONE) ls[rd_idx(rd[s])] = yo[s];
  lds_sync();
  double om[DEB_NOUT_FIELDS];
#pragma unroll
  for (int f = 0; f < DEB_NOUT_FIELDS; ++f) om[f] = w[f];
  {
    TD acc = tconst<TD>(0.0);
    output_fields_t<TD>(cx, [&](int i) { return tconst<TD>(ls[i]); }, L0, L1, L2, P.nq, P.n,
                        [&](int f, const TD &v) { acc = acc + v * om[f]; });
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) thb[j] += acc.d[j + 1];
  }
  if (lane < A.n_in) {
    const int seed = A.in_idx[lane];
    dual acc;
    acc = 0.0;
    output_fields_t<dual>(cx,
                          [&](int i) {
                            dual r;
                            r.v = ls[i];
                            r.d = (i == seed) ? 1.0 : 0.0;
                            return r;
                          },
                          L0, L1, L2, P.nq, P.n, [&](int f, const dual &v) { acc = acc + v * om[f]; });
    lg[seed] = acc.d;
  }
  lds_sync();
#pragma unroll
  for (int s = 0; s < NS; ++s) yb[s] = rd_type(rd[s]) != R_NONE ? lg[rd_idx(rd[s])] : 0.0;
  lds_sync();
}

// one (cosmology, k) item of the gradient kernel.  ring: this wave's ring
// (ring_doubles(NS, C) doubles of device memory, owned for the item's life);
// KBl, ls, lg: per-wave LDS scratch; cx: the wave's context record, in LDS in
// the kernel (its fields are wave-uniform, so one copy serves the wave, and the
// row passes read it there instead of from a per-lane stack copy; every lane
// stores the same values into it)
template <int NS>
__device__ __forceinline__ void adjoint_mode(const Params<NS> &P, const AdjArgs &A, int ci, int ki, double *ring,
                                             double *KBl, double *ls, double *lg, lds_double *pcr, Ctx &cx) {
  const int lane = lane_id();
  const long long item = (long long)ci * P.nk + ki;
  double *gout = A.grad + item * A.ndir;
  if (A.status[item] != DEB_STATUS_OK || A.tape.count[item] < 2) {
    if (lane < A.ndir) gout[lane] = __builtin_nan("");
    return;
  }
  const double k = P.kmodes[item];
  make_ctx<NS>(cx, P, ci, k);
  cx.ndir = A.ndir;
  cx.dcos = A.dcos + (long long)ci * A.ndir * COSMO_DOUBLES;
  cx.dtab = A.dtab + (long long)ci * A.ndir * P.table_stride;
  __shared__ ModeU mu;
  Mode<NS> m;
  m.pcr = pcr;
  m.u = (lds_mode_u *)&mu;
#pragma unroll
  for (int s = 0; s < NS; ++s) m.rd[s] = P.rowdesc[s][lane];
  m.nlev = mode_levels(P.cfg, P.nq);
  const double *tout = P.tau_out + (long long)ci * P.nout;
  const double *dtout = A.dtau + (long long)ci * A.ndir * P.nout;
  int imin = 0, imax = 0;
  for (int i = 1; i < P.nout; ++i) {
    if (tout[i] < tout[imin]) imin = i;
    if (tout[i] > tout[imax]) imax = i;
  }
  // the tape: N accepted steps in segments of C, the last segment's chunk
  // first (a chunk that holds only the final state, when C divides N, is skipped)
  const TapeView &tv = A.tape;
  const int C = tv.ckpt;
  const int N = tv.count[item] - 1;
  auto chunk_t = [&](long long c) { return reinterpret_cast<const double *>(tv.pool + c * tv.chunk_bytes + 64); };
  auto chunk_y = [&](long long c) { return reinterpret_cast<const double *>(tv.pool + c * tv.chunk_bytes + tv.state_off); };
  auto chunk_prev = [&](long long c) { return *reinterpret_cast<const long long *>(tv.pool + c * tv.chunk_bytes); };
  long long chunk = tv.last[item];
  int seg = N / C;
  if (N % C == 0) {
    chunk = chunk_prev(chunk);
    --seg;
  }
  double t0 = 0.0;  // the first state's time
  {
    long long c = chunk;
    for (int s2 = seg; s2 > 0; --s2) c = chunk_prev(c);
    t0 = chunk_t(c)[0];
  }
  auto ring_y = [&](int j) { return ring + (long long)j * NS * WAVE; };
  auto ring_k = [&](int j) { return ring + ((long long)(C + 1) + 8LL * j) * NS * WAVE; };
  const double T1 = tout[imax];
  const double span = T1 - t0;
  double thb[MAXDIR];
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j) thb[j] = 0.0;
  double Tb0 = 0.0, Tb1 = 0.0;
  double yb[NS], pend[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) yb[s] = 0.0;
  const double *wbase = A.ybar + item * (long long)P.nout * DEB_NOUT_FIELDS;
  int iout = P.nout - 1;
  for (; seg >= 0; --seg, chunk = chunk_prev(chunk)) {
    const int nL = min(C, N - seg * C);
    const double *tt = chunk_t(chunk);
    // forward again over the segment from its checkpoint: states and stage
    // vectors into the ring
    {
      double y[NS], y1[NS];
      const double *yc = chunk_y(chunk);
#pragma unroll
      for (int s = 0; s < NS; ++s) y[s] = yc[s * WAVE + lane];
      for (int j = 0; j < nL; ++j) {
        double *ry = ring_y(j);
#pragma unroll
        for (int s = 0; s < NS; ++s) ry[s * WAVE + lane] = y[s];
        step_store<NS>(m, cx, tt[j], tt[j + 1] - tt[j], y, y1, ring_k(j), (lds_double *)KBl);
#pragma unroll
        for (int s = 0; s < NS; ++s) y[s] = y1[s];
      }
      double *ry = ring_y(nL);
#pragma unroll
      for (int s = 0; s < NS; ++s) ry[s * WAVE + lane] = y[s];
    }
    for (int j = nL - 1; j >= 0; --j) {
      const int n = seg * C + j;
      double ycur[NS], ynext[NS];
      {
        const double *r0 = ring_y(j), *r1 = ring_y(j + 1);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          ycur[s] = r0[s * WAVE + lane];
          ynext[s] = r1[s * WAVE + lane];
        }
      }
      const double tcur = tt[j];
      const double h = tt[j + 1] - tcur;
      double tb = 0.0, hb = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) pend[s] = 0.0;
      // outputs interpolated inside this step (t_n < tau_o <= t_{n+1}; step 0
      // also takes tau_o <= t_1)
      while (iout >= 0 && (tout[iout] > tcur || n == 0)) {
        const double th = (tout[iout] - tcur) / h;
        double yo[NS], ybo[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) yo[s] = ycur[s] + th * (ynext[s] - ycur[s]);
        output_vjp<NS>(P, cx, A, m.rd, yo, wbase + (long long)iout * DEB_NOUT_FIELDS, ls, lg, ybo, thb);
        double pth = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          yb[s] += th * ybo[s];
          pend[s] += (1.0 - th) * ybo[s];
          pth += ybo[s] * (ynext[s] - ycur[s]);
        }
        const double thbar = wave_sum1(pth);
        const double taub = thbar / h;
#pragma unroll
        for (int jj = 0; jj < MAXDIR; ++jj)
          if (jj < A.ndir) thb[jj] += taub * dtout[jj * P.nout + iout];
        tb += -thbar / h;
        hb += -thbar * th / h;
        --iout;
      }
      rodas5_step_vjp<NS>(m, cx, tcur, h, ycur, yb, thb, tb, hb, ring_k(j), KBl);
#pragma unroll
      for (int s = 0; s < NS; ++s) yb[s] += pend[s];
      const double sn = (tcur - t0) / span, dsn = h / span;
      Tb0 += tb * (1.0 - sn) - hb * dsn;
      Tb1 += tb * sn + hb * dsn;
    }
  }
  // initial conditions y(t0) = IC(t0)
  {
    TD y0t[NS];
    ics_t<NS, TD>(cx, m.rd, seed_td(t0, 1.0), y0t);
    TD acc[1];
    acc[0] = tconst<TD>(0.0);
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[0] = acc[0] + y0t[s] * yb[s];
    t_wave_sum<TD, 1>(acc);
    Tb0 += acc[0].d[0];
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) thb[j] += acc[0].d[j + 1];
  }
  // t0 = start time(min tau_out), t1 = max tau_out
  {
    TD tmin;
    tmin = tout[imin];
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) tmin.d[j + 1] = j < A.ndir ? dtout[j * P.nout + imin] : 0.0;
    TD ts = start_time_t<TD>(cx, tmin);
#pragma unroll
    for (int j = 0; j < MAXDIR; ++j) {
      thb[j] += Tb0 * ts.d[j + 1];
      if (j < A.ndir) thb[j] += Tb1 * dtout[j * P.nout + imax];
    }
  }
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j)
    if (lane == j && j < A.ndir) gout[j] = thb[j];
}

}  // namespace deb
