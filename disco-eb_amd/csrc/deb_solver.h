// Structured linear algebra and the Rodas5 step for one wave-resident mode.
//
// W = I/(gamma h) - J with J split as
//   J = [ g'(a)    0   0    ]   rows/cols: a (slot 0 lane 0),
//       [ J_1a     0   c_1^T]              a*h' accumulator (lane 1),
//       [ J_Ra     0   B + U C^T ]         the rest (R)
// B is block-tridiagonal along the lane chains (plus the delta_gamma leaf),
// U C^T is the rank-3 coupling through h', eta' and Pi.  Solve: k_a first
// (scalar), then W_B^{-1} by parallel cyclic reduction with a Woodbury
// correction, then k_1 (a*h' is never read by any other row).
#pragma once

#include "deb_mode.h"

namespace deb {

// wave-uniform step factors of Mode
struct ModeU {
  double gh;         // gamma * h
  double inv_kaden;  // 1 / (1/(gamma h) - g')
  double J1a;
  double leafm, wdg, ddg;  // leaf elimination: W_tg,dg / d_dg, W_dg,tg, 1/d_dg
  double Sinv[3][3];
  double c1Z[3];
};
typedef DEB_LDS ModeU lds_mode_u;

template <int NS>
struct Mode {
  uint32_t rd[NS];
  // PCR factors of W_B: d' = sig*d + al*d[i-s] + be*d[i+s] at level s = 2^L;
  // al and be live in the wave's LDS ([PCR_LEVELS][2][NS][64], pcr_lds_doubles)
  lds_double *pcr;
  __device__ __forceinline__ lds_double &al(int L, int j) const { return pcr[((L * 2) * NS + j) * WAVE + lane_id()]; }
  __device__ __forceinline__ lds_double &be(int L, int j) const {
    return pcr[((L * 2 + 1) * NS + j) * WAVE + lane_id()];
  }
  double sig0[NS];
  double binv[NS];
  double Z[3][NS];   // W_B^{-1} [u_h, u_eta, u_pi]
  double Cf[3][NS];  // c_h, c_eta, c_pi
  double c1[NS];
  double U[3][NS];   // u_h, u_eta, u_pi (where h', eta', Pi enter each row)
  double jcol[NS];   // df/da at the step start
  // the wave-uniform factors, in the wave's LDS (one copy per wave rather than
  // one per lane in registers or scratch)
  lds_mode_u *u;
  int nlev;
};

// evaluate f (T-valued) at (tau, y); optionally return the coefficient rows
template <int NS, typename T>
__device__ __forceinline__ void eval_rhs(Ctx &cx, const uint32_t (&rd)[NS], double tau, const double (&y)[NS],
                                         T (&f)[NS], Coef<T> (*co)[NS] = nullptr, T *aH_out = nullptr) {
  const int lane = lane_id();
  double a0 = readlane(y[0], L_A);
  T a;
  if constexpr (sizeof(T) == sizeof(double)) a = a0; else a = T{a0, 1.0};
  Bg<T> b = background<T>(cx, a);
  const double invtau = 1.0 / tau;
  Coef<T> cf[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    // five slots (lmax 50) keep the switch: there the stage vectors stay in
    // registers, and the selects' extra live values cost more than the
    // branches they save
    if constexpr (NS <= 4) cf[s] = row_coef_slot<T>(rd[s], b, cx, invtau, s == 0);
    else cf[s] = row_coef<T>(rd[s], b, cx, invtau);
  }
  // functionals h', eta', Pi, f_1
  constexpr int NV = (sizeof(T) == sizeof(double)) ? 4 : 8;
  double part[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) part[i] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    part[0] += vval(cf[s].ch) * y[s];
    part[1] += vval(cf[s].ceta) * y[s];
    part[2] += vval(cf[s].cpi) * y[s];
    part[3] += vval(cf[s].c1) * y[s];
    if constexpr (NV == 8) {
      part[4] += dval(cf[s].ch) * y[s];
      part[5] += dval(cf[s].ceta) * y[s];
      part[6] += dval(cf[s].cpi) * y[s];
      part[7] += dval(cf[s].c1) * y[s];
    }
  }
  wave_sum<NV>(part);
  T hp, etap, Pi, f1;
  if constexpr (NV == 8) {
    hp = T{part[0], part[4]};
    etap = T{part[1], part[5]};
    Pi = T{part[2], part[6]};
    f1 = T{part[3], part[7]};
  } else {
    hp = part[0];
    etap = part[1];
    Pi = part[2];
    f1 = part[3];
  }
  const double ydg = readlane(y[0], L_DG), ytg = readlane(y[0], L_TG);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double ym = from_prev(y[s]);
    double yp = from_next(y[s]);
    const Coef<T> &r = cf[s];
    T v = r.sub * ym + r.diag * y[s] + r.sup * yp + r.uh * hp + r.ueta * etap + r.upi * Pi;
    if (s == 0) {
      if (lane == L_DG) v = v + r.leaf * ytg;
      if (lane == L_TG) v = v + r.leaf * ydg;
      if (lane == L_A) v = b.aH * b.a;
      if (lane == L_AHP) v = f1;
    }
    f[s] = v;
  }
  if (co) {
#pragma unroll
    for (int s = 0; s < NS; ++s) (*co)[s] = cf[s];
  }
  if (aH_out) *aH_out = b.aH;
}

// df/dtau: only the hierarchy truncation rows depend on tau explicitly
template <int NS>
__device__ __forceinline__ void dfdtau(const uint32_t (&rd)[NS], double tau, const double (&y)[NS], double (&dT)[NS]) {
  const double it = 1.0 / tau, it2 = it * it;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    uint32_t t = rd_type(rd[s]);
    bool trunc = (t == R_FLMAX) | (t == R_GLMAX) | (t == R_FRLMAX) | (t == R_QLMAX);
    dT[s] = trunc ? (rd_l(rd[s]) + 1.0) * it2 * y[s] : 0.0;
  }
}

// W_B^{-1} r (leaf elimination + PCR)
template <int NS>
__device__ __forceinline__ void bsolve(const Mode<NS> &m, double (&x)[NS]) {
  const int lane = lane_id();
  const double rdg = readlane(x[0], L_DG);
  if (lane == L_TG) x[0] -= m.u->leafm * rdg;
#pragma unroll
  for (int L = 0; L < PCR_LEVELS; ++L) {
    if (L >= m.nlev) break;
    const int s = 1 << L;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      double dm, dp;
      if (L == 0) {
        dm = from_prev(x[j]);
        dp = from_next(x[j]);
      } else {
        int lm = lane - s < 0 ? 0 : lane - s;
        int lp = lane + s > WAVE - 1 ? WAVE - 1 : lane + s;
        dm = from_lane(x[j], lm);
        dp = from_lane(x[j], lp);
      }
      double sg = (L == 0) ? m.sig0[j] : 1.0;
      x[j] = sg * x[j] + m.al(L, j) * dm + m.be(L, j) * dp;
    }
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) x[j] *= m.binv[j];
  const double xtg = readlane(x[0], L_TG);
  if (lane == L_DG) x[0] = (rdg - m.u->wdg * xtg) * m.u->ddg;
}

// factorisation of W at (tau0, y0) for step size h
template <int NS>
__device__ __forceinline__ void factor(Mode<NS> &m, Ctx &cx, double tau0, const double (&y0)[NS], double h) {
  const int lane = lane_id();
  constexpr double gamma = 0.19;
  dual fd[NS];
  Coef<dual> co[NS];
  eval_rhs<NS, dual>(cx, m.rd, tau0, y0, fd, &co);
  m.u->gh = gamma * h;
  const double igh = 1.0 / m.u->gh;
  double gprime = readlane(fd[0].d, L_A);
  m.u->J1a = readlane(fd[0].d, L_AHP);
  m.u->inv_kaden = 1.0 / (igh - gprime);
  double A[NS], Bd[NS], Cc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    uint32_t t = rd_type(m.rd[s]);
    bool special = (t == R_NONE) | (t == R_A) | (t == R_AHP);
    m.jcol[s] = special ? 0.0 : fd[s].d;
    A[s] = special ? 0.0 : -co[s].sub.v;
    Cc[s] = special ? 0.0 : -co[s].sup.v;
    Bd[s] = special ? 1.0 : igh - co[s].diag.v;
    m.Cf[0][s] = co[s].ch.v;
    m.Cf[1][s] = co[s].ceta.v;
    m.Cf[2][s] = co[s].cpi.v;
    m.c1[s] = co[s].c1.v;
    m.U[0][s] = co[s].uh.v;
    m.U[1][s] = co[s].ueta.v;
    m.U[2][s] = co[s].upi.v;
  }
  // leaf: delta_gamma (lane L_DG) couples only to theta_gamma (lane L_TG)
  const double w_dg_tg = -readlane(co[0].leaf.v, L_DG);
  const double w_tg_dg = -readlane(co[0].leaf.v, L_TG);
  m.u->ddg = m.u->gh;  // 1 / d_dg with d_dg = 1/(gamma h)
  m.u->wdg = w_dg_tg;
  m.u->leafm = w_tg_dg * m.u->gh;
  if (lane == L_TG) Bd[0] -= w_tg_dg * w_dg_tg * m.u->gh;
  // PCR factorisation; the dark-energy 2x2 block (lanes L_DQ, L_TQ of slot 0)
  // is eliminated at level 0 by cross-multiplication (its theta_Q pivot can
  // vanish), every other row by the usual reduction, which multiplies by the
  // neighbours' pivot reciprocals (each lane takes the reciprocal of its own
  // pivot once per level; rows not coupled to a neighbour never use it)
#pragma unroll
  for (int L = 0; L < PCR_LEVELS; ++L) {
    if (L >= m.nlev) break;
    const int s = 1 << L;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double ib = rcp_nr(Bd[j]);
      double am, ibm, cm, ap, ibp, cp;
      if (L == 0) {
        am = from_prev(A[j]); ibm = from_prev(ib); cm = from_prev(Cc[j]);
        ap = from_next(A[j]); ibp = from_next(ib); cp = from_next(Cc[j]);
      } else {
        int lm = lane - s < 0 ? 0 : lane - s;
        int lp = lane + s > WAVE - 1 ? WAVE - 1 : lane + s;
        am = from_lane(A[j], lm); ibm = from_lane(ib, lm); cm = from_lane(Cc[j], lm);
        ap = from_lane(A[j], lp); ibp = from_lane(ib, lp); cp = from_lane(Cc[j], lp);
      }
      double alpha = (A[j] == 0.0) ? 0.0 : -A[j] * ibm;
      double beta = (Cc[j] == 0.0) ? 0.0 : -Cc[j] * ibp;
      double sg = 1.0;
      if (L == 0 && j == 0) {
        // scale the row by the partner's pivot instead of dividing by it
        const double bm = from_prev(Bd[0]), bp = from_next(Bd[0]);
        if (lane == L_DQ) { sg = bp; beta = -Cc[j]; alpha = 0.0; }
        if (lane == L_TQ) { sg = bm; alpha = -A[j]; beta = 0.0; }
      }
      double nb = sg * Bd[j] + alpha * cm + beta * ap;
      double na = alpha * am;
      double nc = beta * cp;
      m.al(L, j) = alpha;
      m.be(L, j) = beta;
      if (L == 0) m.sig0[j] = sg;
      Bd[j] = nb;
      A[j] = na;
      Cc[j] = nc;
    }
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) m.binv[j] = rcp_nr(Bd[j]);
  // Woodbury: Z = W_B^{-1} U, S = I - C^T Z
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double x[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      uint32_t t = rd_type(m.rd[s]);
      bool special = (t == R_NONE) | (t == R_A) | (t == R_AHP);
      double u = r == 0 ? co[s].uh.v : (r == 1 ? co[s].ueta.v : co[s].upi.v);
      x[s] = special ? 0.0 : u;
    }
    bsolve<NS>(m, x);
#pragma unroll
    for (int s = 0; s < NS; ++s) m.Z[r][s] = x[s];
  }
  double part[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) part[i] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) part[r * 3 + c] += m.Cf[r][s] * m.Z[c][s];
      part[9 + r] += m.c1[s] * m.Z[r][s];
    }
  }
  wave_sum<12>(part);
  double S[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[r][c] = (r == c ? 1.0 : 0.0) - part[r * 3 + c];
  const double det = S[0][0] * (S[1][1] * S[2][2] - S[1][2] * S[2][1]) -
                     S[0][1] * (S[1][0] * S[2][2] - S[1][2] * S[2][0]) +
                     S[0][2] * (S[1][0] * S[2][1] - S[1][1] * S[2][0]);
  const double id = 1.0 / det;
  m.u->Sinv[0][0] = (S[1][1] * S[2][2] - S[1][2] * S[2][1]) * id;
  m.u->Sinv[0][1] = (S[0][2] * S[2][1] - S[0][1] * S[2][2]) * id;
  m.u->Sinv[0][2] = (S[0][1] * S[1][2] - S[0][2] * S[1][1]) * id;
  m.u->Sinv[1][0] = (S[1][2] * S[2][0] - S[1][0] * S[2][2]) * id;
  m.u->Sinv[1][1] = (S[0][0] * S[2][2] - S[0][2] * S[2][0]) * id;
  m.u->Sinv[1][2] = (S[0][2] * S[1][0] - S[0][0] * S[1][2]) * id;
  m.u->Sinv[2][0] = (S[1][0] * S[2][1] - S[1][1] * S[2][0]) * id;
  m.u->Sinv[2][1] = (S[0][1] * S[2][0] - S[0][0] * S[2][1]) * id;
  m.u->Sinv[2][2] = (S[0][0] * S[1][1] - S[0][1] * S[1][0]) * id;
  m.u->c1Z[0] = part[9];
  m.u->c1Z[1] = part[10];
  m.u->c1Z[2] = part[11];
}

// k = W^{-1} r (r is overwritten by k)
template <int NS>
__device__ __forceinline__ void wsolve(const Mode<NS> &m, double (&r)[NS]) {
  const int lane = lane_id();
  const double ra = readlane(r[0], L_A);
  const double r1 = readlane(r[0], L_AHP);
  const double ka = ra * m.u->inv_kaden;
#pragma unroll
  for (int s = 0; s < NS; ++s) r[s] += m.jcol[s] * ka;
  if (lane == L_A || lane == L_AHP) r[0] = 0.0;
  bsolve<NS>(m, r);
  double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    g[0] += m.Cf[0][s] * r[s];
    g[1] += m.Cf[1][s] * r[s];
    g[2] += m.Cf[2][s] * r[s];
    g[3] += m.c1[s] * r[s];
  }
  wave_sum<4>(g);
  double t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = m.u->Sinv[i][0] * g[0] + m.u->Sinv[i][1] * g[1] + m.u->Sinv[i][2] * g[2];
#pragma unroll
  for (int s = 0; s < NS; ++s) r[s] += m.Z[0][s] * t[0] + m.Z[1][s] * t[1] + m.Z[2][s] * t[2];
  const double c1k = g[3] + m.u->c1Z[0] * t[0] + m.u->c1Z[1] * t[1] + m.u->c1Z[2] * t[2];
  const double k1 = m.u->gh * (r1 + m.u->J1a * ka + c1k);
  if (lane == L_A) r[0] = ka;
  if (lane == L_AHP) r[0] = k1;
}

// Rodas5 (W-form) tableau: ode_integrators_stiff.py:622-687
struct R5 {
  static constexpr double gamma = 0.19;
  static constexpr double a21 = 2.0, a31 = 3.040894194418781, a32 = 1.041747909077569, a41 = 2.576417536461461,
                          a42 = 1.622083060776640, a43 = -0.9089668560264532, a51 = 2.760842080225597,
                          a52 = 1.446624659844071, a53 = -0.3036980084553738, a54 = 0.2877498600325443,
                          a61 = -14.09640773051259, a62 = 6.925207756232704, a63 = -41.47510893210728,
                          a64 = 2.343771018586405, a65 = 24.13215229196062;
  static constexpr double C21 = -10.31323885133993, C31 = -21.04823117650003, C32 = -7.234992135176716,
                          C41 = 32.22751541853323, C42 = -4.943732386540191, C43 = 19.44922031041879,
                          C51 = -20.69865579590063, C52 = -8.816374604402768, C53 = 1.260436877740897,
                          C54 = -0.7495647613787146, C61 = -46.22004352711257, C62 = -17.49534862857472,
                          C63 = -289.6389582892057, C64 = 93.60855400400906, C65 = 318.3822534212147,
                          C71 = 34.20013733472935, C72 = -14.15535402717690, C73 = 57.82335640988400,
                          C74 = 25.83362985412365, C75 = 1.408950972071624, C76 = -6.551835421242162,
                          C81 = 42.57076742291101, C82 = -13.80770672017997, C83 = 93.98938432427124,
                          C84 = 18.77919633714503, C85 = -31.58359187223370, C86 = -6.685968952921985,
                          C87 = -5.810979938412932;
  static constexpr double c2 = 0.38, c3 = 0.3878509998321533, c4 = 0.4839718937873840, c5 = 0.4570477008819580;
  static constexpr double d1 = 0.19, d2 = -0.1823079225333714636, d3 = -0.319231832186874912,
                          d4 = 0.3449828624725343, d5 = -0.377417564392089818;
};

// Out-of-line versions of the factorisation and the right-hand side, for
// kernels that run the step next to much other live state (the reverse
// kernel): as calls, the step factors stay in memory instead of registers.
template <int NS>
__device__ __noinline__ void factor_call(Mode<NS> &m, Ctx &cx, double tau0, const double (&y0)[NS], double h) {
  factor<NS>(m, cx, tau0, y0, h);
}
// (the calls work on private copies of their array arguments: stores through
// the references could alias the inputs as far as the compiler knows, which
// would reload them from the caller's frame after every store)
template <int NS>
__device__ __noinline__ void rhs_call(Ctx &cx, const uint32_t (&rd)[NS], double tau, const double (&y_in)[NS],
                                      double (&f_out)[NS]) {
  double y[NS], f[NS];
  uint32_t d[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    y[s] = y_in[s];
    d[s] = rd[s];
  }
  eval_rhs<NS, double>(cx, d, tau, y, f);
#pragma unroll
  for (int s = 0; s < NS; ++s) f_out[s] = f[s];
}

// one Rodas5Transformed step (ode_integrators_stiff.py:718-834): y1 and err = k8;
// st(i, k) sees every stage vector k_{i+1} once it is solved.  CALLS selects
// the out-of-line factorisation and right-hand side.  With kl (the
// wave's LDS, 7 x NS x 64 doubles) the stage vectors k1..k7 are kept there
// instead of in registers, which relieves the register file of the kernels
// that run the step fully inlined.
template <int NS, class Store, bool CALLS = false>
__device__ __forceinline__ void rodas5_step_s(Mode<NS> &m, Ctx &cx, double t0, double dt, const double (&y0)[NS],
                                              double (&y1)[NS], double (&err)[NS], Store st,
                                              lds_double *kl = nullptr) {
  using R = R5;
  const int lane = lane_id();
  auto eval_rhs_ = [&](double tau, const double (&y)[NS], double (&f)[NS]) {
    if constexpr (CALLS) rhs_call<NS>(cx, m.rd, tau, y, f);
    else eval_rhs<NS, double>(cx, m.rd, tau, y, f);
  };
  // the solves stay inline in both forms (as a call each, the recompute of the
  // reverse kernel measured 2 % slower)
  auto wsolve_ = [&](double (&r)[NS]) { wsolve<NS>(m, r); };
  if constexpr (CALLS) factor_call<NS>(m, cx, t0, y0, dt);
  else factor<NS>(m, cx, t0, y0, dt);
  double dT[NS];
  dfdtau<NS>(m.rd, t0, y0, dT);
  double kr[7][NS];
  // stage vector i (0-based) of slot s: registers, or LDS when kl is given
  auto K = [&](int i, int s) -> double { return kl ? (double)kl[(i * NS + s) * WAVE + lane] : kr[i][s]; };
  auto keep = [&](int i, const double (&k)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (kl) kl[(i * NS + s) * WAVE + lane] = k[s];
      else kr[i][s] = k[s];
    }
  };
  double k[NS], u[NS], f[NS];
  const double idt = 1.0 / dt;
  // stage 1
  eval_rhs_(t0, y0, f);
#pragma unroll
  for (int s = 0; s < NS; ++s) k[s] = f[s] + dt * R::d1 * dT[s];
  wsolve_(k);
  st(0, k);
  keep(0, k);
  // stage 2
#pragma unroll
  for (int s = 0; s < NS; ++s) u[s] = y0[s] + R::a21 * K(0, s);
  eval_rhs_(t0 + R::c2 * dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s) k[s] = f[s] + dt * R::d2 * dT[s] + (R::C21 * idt) * K(0, s);
  wsolve_(k);
  st(1, k);
  keep(1, k);
  // stage 3
#pragma unroll
  for (int s = 0; s < NS; ++s) u[s] = y0[s] + R::a31 * K(0, s) + R::a32 * K(1, s);
  eval_rhs_(t0 + R::c3 * dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    k[s] = f[s] + dt * R::d3 * dT[s] + ((R::C31 * idt) * K(0, s) + (R::C32 * idt) * K(1, s));
  wsolve_(k);
  st(2, k);
  keep(2, k);
  // stage 4
#pragma unroll
  for (int s = 0; s < NS; ++s) u[s] = y0[s] + R::a41 * K(0, s) + R::a42 * K(1, s) + R::a43 * K(2, s);
  eval_rhs_(t0 + R::c4 * dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    k[s] = f[s] + dt * R::d4 * dT[s] +
           ((R::C41 * idt) * K(0, s) + (R::C42 * idt) * K(1, s) + (R::C43 * idt) * K(2, s));
  wsolve_(k);
  st(3, k);
  keep(3, k);
  // stage 5
#pragma unroll
  for (int s = 0; s < NS; ++s)
    u[s] = y0[s] + R::a51 * K(0, s) + R::a52 * K(1, s) + R::a53 * K(2, s) + R::a54 * K(3, s);
  eval_rhs_(t0 + R::c5 * dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    k[s] = f[s] + dt * R::d5 * dT[s] +
           ((R::C51 * idt) * K(0, s) + (R::C52 * idt) * K(1, s) + (R::C53 * idt) * K(2, s) +
            (R::C54 * idt) * K(3, s));
  wsolve_(k);
  st(4, k);
  keep(4, k);
  // stage 6
#pragma unroll
  for (int s = 0; s < NS; ++s)
    u[s] = y0[s] + R::a61 * K(0, s) + R::a62 * K(1, s) + R::a63 * K(2, s) + R::a64 * K(3, s) + R::a65 * K(4, s);
  eval_rhs_(t0 + dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    k[s] = f[s] + ((R::C61 * idt) * K(0, s) + (R::C62 * idt) * K(1, s) + (R::C63 * idt) * K(2, s) +
                   (R::C64 * idt) * K(3, s) + (R::C65 * idt) * K(4, s));
  wsolve_(k);
  st(5, k);
  keep(5, k);
  // stage 7
#pragma unroll
  for (int s = 0; s < NS; ++s) u[s] = u[s] + k[s];
  eval_rhs_(t0 + dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    k[s] = f[s] + ((R::C71 * idt) * K(0, s) + (R::C72 * idt) * K(1, s) + (R::C73 * idt) * K(2, s) +
                   (R::C74 * idt) * K(3, s) + (R::C75 * idt) * K(4, s) + (R::C76 * idt) * K(5, s));
  wsolve_(k);
  st(6, k);
  keep(6, k);
  // stage 8
#pragma unroll
  for (int s = 0; s < NS; ++s) u[s] = u[s] + k[s];
  eval_rhs_(t0 + dt, u, f);
#pragma unroll
  for (int s = 0; s < NS; ++s)
    err[s] = f[s] + ((R::C81 * idt) * K(0, s) + (R::C82 * idt) * K(1, s) + (R::C83 * idt) * K(2, s) +
                     (R::C84 * idt) * K(3, s) + (R::C85 * idt) * K(4, s) + (R::C86 * idt) * K(5, s) +
                     (R::C87 * idt) * K(6, s));
  wsolve_(err);
  st(7, err);
#pragma unroll
  for (int s = 0; s < NS; ++s) y1[s] = u[s] + err[s];
}

template <int NS>
__device__ __forceinline__ void rodas5_step(Mode<NS> &m, Ctx &cx, double t0, double dt, const double (&y0)[NS],
                                            double (&y1)[NS], double (&err)[NS], lds_double *kl = nullptr) {
  rodas5_step_s<NS>(m, cx, t0, dt, y0, y1, err, [](int, const double (&)[NS]) {}, kl);
}

}  // namespace deb
