// gfx950 kernels and C ABI of the DISCO-EB perturbation hot path.
//
//   deb_forward_kernel  one wavefront per (cosmology, k) item (largest k
//                       dispatched first), runs
//                       determine_starting_time, adiabatic ICs and the
//                       adaptive Rodas5 loop (diffrax PID semantics), and
//                       writes the output fields at every tau_out.
//   deb_rhs_kernel      test hook: model_synchronous for given states.
//   deb_step_kernel     test hook: one Rodas5Transformed step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "deb_solver.h"

namespace deb {

// ---------------------------------------------------------------------------
// per-mode setup: determine_starting_time (perturbations.py:630-681) and
// adiabatic_ics_one_mode (perturbations.py:526-627)

__device__ __forceinline__ double start_time(Ctx &cx, double tau_out_min) {
  const deb_cosmo_t &c = *cx.c;
  const double tau0 = c.taumin;
  const double tau1 = spl_eval(cx.sp[DEB_SPL_TAU_A], 0.1, cx.hint[DEB_SPL_TAU_A]);
  const double tau_k = 1.0 / cx.k;
  const double akthom = 2.3048e-9 * (1.0 - c.YHe) * c.Omegab * c.H0 * c.H0;
  double lt[2];
  for (int which = 0; which < 2; ++which) {
    double xl = log(tau0), xr = log(tau1);
    auto fn = [&](double x) {
      double tau = exp(x);
      double a = spl_eval(cx.sp[DEB_SPL_A_TAU], tau, cx.hint[DEB_SPL_A_TAU]);
      Bg<double> b = background<double>(cx, a);
      if (which == 0) return (1.0 / b.aH) / tau_k / 0.07 - 1.0;
      double xe = spl_eval(cx.sp[DEB_SPL_XE_TAU], tau, cx.hint[DEB_SPL_XE_TAU]);
      double opac = xe * akthom / (a * a);
      return (1.0 / opac) / (1.0 / b.aH) / 0.0004 - 1.0;
    };
    for (int i = 0; i < 7; ++i) {
      double xm = 0.5 * (xl + xr);
      if (fn(xm) * fn(xl) > 0) xl = xm; else xr = xm;
    }
    lt[which] = 0.5 * (xl + xr);
  }
  double ts = exp(fmin(lt[1], lt[0]));
  return 0.99 * fmin(tau_out_min, ts);
}

template <int NS>
__device__ __forceinline__ void initial_conditions(Ctx &cx, const uint32_t (&rd)[NS], double tau, double (&y)[NS]) {
  const deb_cosmo_t &c = *cx.c;
  const double k = cx.k;
  double a = spl_eval(cx.sp[DEB_SPL_A_TAU], tau, cx.hint[DEB_SPL_A_TAU]);
  double rnu = exp(spl_eval(cx.sp[DEB_SPL_LOGRHONU_LOGA], log(a), cx.hint[DEB_SPL_LOGRHONU_LOGA]));
  double a4 = a * a * a * a;
  double rhom = c.grhom * c.Omegam / (a * a * a);
  double rhor = (c.grhog + c.grhor * (c.Neff + c.Nmnu * rnu)) / a4;
  double rhonu = c.grhor * (c.Neff + c.Nmnu * rnu) / a4;
  double fracb = c.Omegab / c.Omegam;
  double fracnu = rhonu / rhor;
  double om = a * rhom / sqrt(rhor);
  const double C = -1.0;
  double kt = k * tau;
  double deltag = -(kt * kt) / 3.0 * (1.0 - om * tau / 5.0) * C;
  double thetag = -(kt * kt * kt) / tau / 36.0 *
                  (1.0 - 3.0 * (1.0 + 5.0 * fracb - fracnu) / 20.0 / (1.0 - fracnu) * om * tau) * C;
  double thetar = -(kt * kt * kt * kt) / tau / 36.0 / (4.0 * fracnu + 15.0) *
                  (4.0 * fracnu + 11.0 + 12.0 -
                   3.0 * (8.0 * fracnu * fracnu + 50.0 * fracnu + 275.0) / 20.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  double shearr = (kt * kt) / (45.0 + 12.0 * fracnu) * 2.0 *
                  (1.0 + (4.0 * fracnu - 5.0) / 4.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  double cs2Q = c.cs2_DE;
  double wQ = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  double deltaq = (kt * kt) / 4.0 * (1.0 + wQ) * (4.0 - 3.0 * cs2Q) / (4.0 - 6.0 * wQ + 3.0 * cs2Q) * C;
  double thetaq = (kt * kt * kt * kt) / tau / 4.0 * cs2Q / (4.0 - 6.0 * wQ + 3.0 * cs2Q) * C;
  if (cx.lambda) {
    deltaq = 0.0;
    thetaq = 0.0;
  }
  double eta = C * (1.0 - (kt * kt) / 12.0 / (15.0 + 4.0 * fracnu) *
                              (5.0 + 4.0 * fracnu - (16.0 * fracnu * fracnu + 280.0 * fracnu + 325.0) / 10.0 /
                                                        (2.0 * fracnu + 15.0) * tau * om));
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    uint32_t t = rd_type(rd[s]);
    int qi = (int)rd_q(rd[s]);
    double v = 0.0;
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
        double aq = a * c.amnu / cx.q[qi];
        double vq = 1.0 / sqrt(1.0 + aq * aq);
        double dl = cx.dlnf[qi];
        if (t == R_Q0) v = -0.25 * dl * deltag;
        else if (t == R_Q1) v = -dl * thetar / vq / k / 3.0;
        else v = -0.5 * dl * shearr;
        break;
      }
      default: v = 0.0; break;
    }
    y[s] = v;
  }
}

// convert_to_output_variables (perturbations.py:374-523) from a state staged
// in LDS (reference index order)
__device__ __forceinline__ void output_fields(Ctx &cx, const double *ys, int lg, int lgp, int lr, int nq, int n, double *o) {
  const deb_cosmo_t &c = *cx.c;
  const double k = cx.k, k2 = k * k;
  const double Omegac = c.Omegam - c.Omegab;
  const int iq0 = 10 + lg + lgp + lr, iq1 = iq0 + nq;
  double a = ys[0], eta = ys[2], deltac = ys[3], thetac = ys[4], deltab = ys[5], thetab = ys[6];
  double deltag = ys[7], thetag = ys[8];
  double deltar = ys[9 + lg + lgp], thetar = ys[10 + lg + lgp];
  double la = log(a);
  double rhonu = exp(spl_eval(cx.sp[DEB_SPL_LOGRHONU_LOGA], la, cx.hint[DEB_SPL_LOGRHONU_LOGA]));
  double pnu = exp(spl_eval(cx.sp[DEB_SPL_LOGPNU_LOGA], la, cx.hint[DEB_SPL_LOGPNU_LOGA]));
  double drhonu = 0.0, fnu = 0.0;
  for (int i = 0; i < nq; ++i) {
    double aq = a * c.amnu / cx.q[i];
    double v = 1.0 / sqrt(1.0 + aq * aq);
    drhonu += cx.w[i] * ys[iq0 + i] / v;
    fnu += cx.w[i] * ys[iq1 + i];
  }
  double deltanu = drhonu / rhonu;
  double thetanu = k * fnu / (rhonu + pnu);
  double deltaq = ys[n - 2], thetaq = ys[n - 1];
  double wQ = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  double rhoQ = pow(a, -3.0 * (1.0 + c.w_DE_0 + c.w_DE_a)) * exp(3.0 * (a - 1.0) * c.w_DE_a);
  double a2 = a * a;
  double rppQ = (1.0 + wQ) * rhoQ * c.grhom * c.OmegaDE * thetaq * a2;
  double grho = c.grhom * c.Omegam / a + (c.grhog + c.grhor * (c.Neff + c.Nmnu * rhonu)) / a2 +
                c.grhom * c.OmegaDE * rhoQ * a2 + c.grhom * c.Omegak;
  double aH = sqrt(grho / 3.0);
  double dgrho = c.grhom * (Omegac * deltac + c.Omegab * deltab) / a +
                 (c.grhog * deltag + c.grhor * (c.Neff * deltar + c.Nmnu * drhonu)) / a2 +
                 c.grhom * c.OmegaDE * deltaq * rhoQ * a2;
  double dgtheta = c.grhom * (Omegac * thetac + c.Omegab * thetab) / a +
                   4.0 / 3.0 * (c.grhog * thetag + c.Neff * c.grhor * thetar) / a2 + c.Nmnu * c.grhor * k * fnu / a2 +
                   rppQ;
  double hprime = (2.0 * k2 * eta + dgrho) / aH;
  double etaprime = 0.5 * dgtheta / k2;
  double alpha = (hprime + 6.0 * etaprime) / 2.0 / k2;
  double mden = c.grhom * c.Omegam / a + c.grhor * c.Nmnu * rhonu / a2;
  double deltam = (c.grhom * (Omegac * deltac + c.Omegab * deltab) / a + c.grhor * c.Nmnu * drhonu / a2) / mden;
  double thetam =
      (c.grhom * (Omegac * thetac + c.Omegab * thetab) / a + c.Nmnu * c.grhor * k * fnu / a2) / (3.0 * mden);
  double deltabc = (c.grhom * (Omegac * deltac + c.Omegab * deltab) / a) / (c.grhom * c.Omegam / a);
  double thetabc = (c.grhom * (Omegac * thetac + c.Omegab * thetab) / a) / (3.0 * (c.grhom * c.Omegam / a) / a2);
  thetam += alpha * k2;
  thetabc += alpha * k2;
  o[0] = eta; o[1] = etaprime; o[2] = hprime; o[3] = alpha;
  o[4] = deltam; o[5] = thetam / aH; o[6] = deltabc; o[7] = thetabc / aH;
  o[8] = deltac; o[9] = thetac / aH; o[10] = deltab; o[11] = thetab / aH;
  o[12] = deltag; o[13] = thetag / aH; o[14] = deltar; o[15] = thetar / aH;
  o[16] = deltanu; o[17] = thetanu / aH; o[18] = deltaq; o[19] = thetaq / aH;
}

template <int NS>
__device__ __forceinline__ void make_ctx(Ctx &cx, const Params<NS> &P, int ci, double k) {
  cx.c = P.cosmo + ci;
  const double *tb = P.tables + (int64_t)ci * P.table_stride;
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    int off = P.desc[s][0], n = P.desc[s][1];
    cx.sp[s] = {tb + off, tb + off + n, tb + off + 2 * n, n};
    cx.hint[s] = -1;
  }
  cx.k = k;
  cx.nq = P.nq;
  cx.q = P.q;
  cx.w = P.w;
  cx.dlnf = P.dlnf;
  cx.lambda = (cx.c->w_DE_0 == -1.0) && (cx.c->w_DE_a == 0.0);
}

template <int NS>
__device__ __forceinline__ void write_outputs(const Params<NS> &P, Ctx &cx, const uint32_t (&rd)[NS], const double (&ys)[NS],
                              double *lds, double *dst) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (rd_type(rd[s]) != R_NONE) lds[rd_idx(rd[s])] = ys[s];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if (P.cfg.return_full) {
    for (int i = lane; i < P.n; i += WAVE) dst[i] = lds[i];
  } else {
    double o[DEB_NOUT_FIELDS];
    output_fields(cx, lds, P.cfg.lmaxg, P.cfg.lmaxgp, P.cfg.lmaxr, P.nq, P.n, o);
#pragma unroll
    for (int j = 0; j < DEB_NOUT_FIELDS; ++j)
      if (lane == j) dst[j] = o[j];
  }
  __builtin_amdgcn_s_barrier();
}

// evolve_one_mode (perturbations.py:728-781) for one (cosmology, k) item
template <int NS>
__device__ __forceinline__ void integrate_mode(const Params<NS> &P, int ci, int ki, double *lds) {
  const int lane = lane_id();
  Ctx cx;
  const double k = P.kmodes[(int64_t)ci * P.nk + ki];
  make_ctx<NS>(cx, P, ci, k);
  Mode<NS> m;
#pragma unroll
  for (int s = 0; s < NS; ++s) m.rd[s] = P.rowdesc[s][lane];
  {
    int maxlen = P.cfg.lmaxg + 2;
    maxlen = max(maxlen, P.cfg.lmaxgp + 1);
    maxlen = max(maxlen, P.cfg.lmaxr + 1);
    if (P.nq > 0) maxlen = max(maxlen, P.cfg.lmaxnu + 1);
    int nl = 0;
    while ((1 << nl) < maxlen) ++nl;
    m.nlev = nl;
  }
  const double *tout = P.tau_out + (int64_t)ci * P.nout;
  double tmin = tout[0], tmax = tout[0];
  for (int i = 1; i < P.nout; ++i) {
    tmin = fmin(tmin, tout[i]);
    tmax = fmax(tmax, tout[i]);
  }
  const double t0 = start_time(cx, tmin);
  const double t1 = tmax;
  double y[NS], y1[NS], err[NS];
  initial_conditions<NS>(cx, m.rd, t0, y);
  const deb_config_t &cf = P.cfg;
  const double eo = 5.0;
  const double coeff1 = (cf.icoeff + cf.pcoeff + cf.dcoeff) / eo;
  const double coeff2 = -(cf.pcoeff + 2.0 * cf.dcoeff) / eo;
  const double coeff3 = cf.dcoeff / eo;
  // error-norm weights: indices (a, eta, delta_c, delta_b, theta_b, delta_gamma) of slot 0
  double ew = 0.0;
  if (lane == L_A || lane == L_DC || lane == L_DB || lane == L_DG) ew = 1.0;
  if (lane == L_ETA) ew = k * k;
  if (lane == L_TB) ew = 1.0 / (k * k);
  double tprev = t0;
  double tnext = fmin(t0 + fmin(t0 / 4.0, 0.5 * (t1 - t0)), t1);
  double prev_inv = 1.0, prev_prev_inv = 1.0;
  int iout = 0, steps = 0;
  int status = DEB_STATUS_OK;
  const int width = cf.return_full ? P.n : DEB_NOUT_FIELDS;
  double *ydst = P.y_out + ((int64_t)ci * P.nk + ki) * (int64_t)P.nout * width;
  while (tprev < t1) {
    if (steps >= cf.max_steps) {
      status = DEB_STATUS_MAX_STEPS;
      break;
    }
    const double dt = tnext - tprev;
    rodas5_step<NS>(m, cx, tprev, dt, y, y1, err);
    double e2 = 0.0;
    {
      double y0s = y[0], y1s = y1[0];
      double sc = err[0] / (cf.atol + fmax(fabs(y0s), fabs(y1s)) * cf.rtol);
      double x = sc * ew;
      e2 = x * x;
    }
    const double nrm = sqrt(wave_sum1(e2) / 6.0);
    if (!isfinite(nrm)) {
      status = DEB_STATUS_NONFINITE;
      break;
    }
    const bool keep = nrm < 1.0;
    double inv = 1.0 / nrm;
    double f1 = coeff1 == 0.0 ? 1.0 : pow(inv, coeff1);
    double f2 = coeff2 == 0.0 ? 1.0 : pow(prev_inv, coeff2);
    double f3 = coeff3 == 0.0 ? 1.0 : pow(prev_prev_inv, coeff3);
    double fmn = keep ? 1.0 : cf.factormin;
    double factor = fmin(fmax(0.9 * f1 * f2 * f3, fmn), cf.factormax);
    double dtn = dt * factor;
    if (inv == 0.0 || isinf(inv)) inv = 1.0;
    if (keep) {
      while (iout < P.nout && tout[iout] <= tnext) {
        double th = (tout[iout] - tprev) / dt;
        double ys[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) ys[s] = y[s] + th * (y1[s] - y[s]);
        write_outputs<NS>(P, cx, m.rd, ys, lds, ydst + (int64_t)iout * width);
        ++iout;
      }
      prev_prev_inv = prev_inv;
      prev_inv = inv;
#pragma unroll
      for (int s = 0; s < NS; ++s) y[s] = y1[s];
    }
    double nt0 = keep ? tnext : tprev;
    double nt1 = nt0 + dtn;
    if (nt1 > t1 - 1e-10) nt1 = keep ? t1 : nt0 + 0.5 * (t1 - nt0);
    tprev = nt0;
    tnext = nt1;
    ++steps;
  }
  if (lane == 0) {
    P.status[(int64_t)ci * P.nk + ki] = status;
    if (P.nsteps) P.nsteps[(int64_t)ci * P.nk + ki] = steps;
  }
}

// One workgroup (one wavefront) per (cosmology, k) item; the hardware
// dispatcher hands the next item to whichever SIMD frees up first, and items
// are numbered so that the largest k (most steps) start first.
template <int NS>
__global__ void __launch_bounds__(WAVE) deb_forward_kernel(const Params<NS> P) {
  __shared__ double lds[WAVE * NS];
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * P.nk) return;
  const int ki = P.nk - 1 - item / P.n_cosmo;
  const int ci = item % P.n_cosmo;
  integrate_mode<NS>(P, ci, ki, lds);
}

// ---------------------------------------------------------------------------
// test hooks
template <int NS>
struct HookArgs {
  const int32_t *ci;
  const double *kk, *t0, *t1, *y0;
  double *y1, *err;
  int m;
};

template <int NS>
__global__ void __launch_bounds__(WAVE) deb_rhs_kernel(const Params<NS> P, const HookArgs<NS> H) {
  const int item = blockIdx.x;
  if (item >= H.m) return;
  const int lane = lane_id();
  Ctx cx;
  make_ctx<NS>(cx, P, H.ci[item], H.kk[item]);
  uint32_t rd[NS];
  double y[NS], f[NS];
  const double *src = H.y0 + (int64_t)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rd[s] = P.rowdesc[s][lane];
    y[s] = rd_type(rd[s]) != R_NONE ? src[rd_idx(rd[s])] : 0.0;
  }
  eval_rhs<NS, double>(cx, rd, H.t0[item], y, f);
  double *dst = H.y1 + (int64_t)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    if (rd_type(rd[s]) != R_NONE) dst[rd_idx(rd[s])] = f[s];
}

template <int NS>
__global__ void __launch_bounds__(WAVE) deb_step_kernel(const Params<NS> P, const HookArgs<NS> H) {
  const int item = blockIdx.x;
  if (item >= H.m) return;
  const int lane = lane_id();
  Ctx cx;
  make_ctx<NS>(cx, P, H.ci[item], H.kk[item]);
  Mode<NS> m;
  {
    int maxlen = P.cfg.lmaxg + 2;
    maxlen = max(maxlen, P.cfg.lmaxgp + 1);
    maxlen = max(maxlen, P.cfg.lmaxr + 1);
    if (P.nq > 0) maxlen = max(maxlen, P.cfg.lmaxnu + 1);
    int nl = 0;
    while ((1 << nl) < maxlen) ++nl;
    m.nlev = nl;
  }
  double y[NS], y1[NS], err[NS];
  const double *src = H.y0 + (int64_t)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    m.rd[s] = P.rowdesc[s][lane];
    y[s] = rd_type(m.rd[s]) != R_NONE ? src[rd_idx(m.rd[s])] : 0.0;
  }
  const double t0 = H.t0[item];
  rodas5_step<NS>(m, cx, t0, H.t1[item] - t0, y, y1, err);
  double *d1 = H.y1 + (int64_t)item * P.n;
  double *de = H.err + (int64_t)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    if (rd_type(m.rd[s]) != R_NONE) {
      d1[rd_idx(m.rd[s])] = y1[s];
      de[rd_idx(m.rd[s])] = err[s];
    }
}

// ---------------------------------------------------------------------------
// host side: lane layout, argument checks, launches

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct Layout {
  int ns = 0;
  std::vector<uint32_t> rd;  // [ns][64]
};

int build_layout(const deb_config_t &c, Layout &L) {
  const int lg = c.lmaxg, lgp = c.lmaxgp, lr = c.lmaxr, lnu = c.lmaxnu, nq = c.nqmax;
  if (lg < 3 || lgp < 3 || lr < 3 || (nq > 0 && lnu < 3))
    return fail(DEB_E_UNSUPPORTED, "hierarchy truncations must be >= 3");
  if (nq < 0 || nq > DEB_MAXNQ) return fail(DEB_E_UNSUPPORTED, "nqmax outside [0, 16]");
  if (8 + lg + 2 > WAVE) return fail(DEB_E_UNSUPPORTED, "lmaxg > 54 does not fit one wavefront");
  if (lgp + 1 > WAVE || lr + 1 > WAVE || lnu + 1 > WAVE)
    return fail(DEB_E_UNSUPPORTED, "a hierarchy longer than 64 multipoles does not fit one wavefront");
  const int n = 12 + lg + lgp + lr + nq * (lnu + 1);
  const int ig = 7, igp = 8 + lg, ir = 9 + lg + lgp, iq0 = 10 + lg + lgp + lr;
  struct Chain {
    std::vector<uint32_t> rows;
  };
  std::vector<Chain> chains;
  {  // polarization
    Chain ch;
    for (int l = 0; l <= lgp; ++l) {
      uint32_t t = l == 0 ? R_G0 : (l == 2 ? R_G2 : (l == lgp ? R_GLMAX : R_GL));
      ch.rows.push_back(rd_make(t, l, 0, igp + l));
    }
    chains.push_back(ch);
  }
  {  // massless neutrinos
    Chain ch;
    for (int l = 0; l <= lr; ++l) {
      uint32_t t = l == 0 ? R_DR : (l == 1 ? R_TR : (l == 2 ? R_FR2 : (l == lr ? R_FRLMAX : R_FRL)));
      ch.rows.push_back(rd_make(t, l, 0, ir + l));
    }
    chains.push_back(ch);
  }
  for (int i = 0; i < nq; ++i) {  // massive neutrinos, one chain per momentum bin
    Chain ch;
    for (int l = 0; l <= lnu; ++l) {
      uint32_t t = l == 0 ? R_Q0 : (l == 1 ? R_Q1 : (l == 2 ? R_Q2 : (l == lnu ? R_QLMAX : R_QL)));
      ch.rows.push_back(rd_make(t, l, i, iq0 + l * nq + i));
    }
    chains.push_back(ch);
  }
  std::vector<std::vector<uint32_t>> slots(1, std::vector<uint32_t>(WAVE, rd_make(R_NONE, 0, 0, 0)));
  auto &s0 = slots[0];
  s0[L_A] = rd_make(R_A, 0, 0, 0);
  s0[L_AHP] = rd_make(R_AHP, 0, 0, 1);
  s0[L_ETA] = rd_make(R_ETA, 0, 0, 2);
  s0[L_DG] = rd_make(R_DG, 0, 0, ig);
  s0[L_DC] = rd_make(R_DC, 0, 0, 3);
  s0[L_TC] = rd_make(R_TC, 0, 0, 4);
  s0[L_DQ] = rd_make(R_DQ, 0, 0, n - 2);
  s0[L_TQ] = rd_make(R_TQ, 0, 0, n - 1);
  s0[L_DB] = rd_make(R_DB, 0, 0, 5);
  s0[L_TB] = rd_make(R_TB, 1, 0, 6);
  s0[L_TG] = rd_make(R_TG, 1, 0, ig + 1);
  for (int l = 2; l <= lg; ++l) {
    uint32_t t = l == 2 ? R_F2 : (l == lg ? R_FLMAX : R_FL);
    s0[L_TG + l - 1] = rd_make(t, l, 0, ig + l);
  }
  std::vector<int> used(1, L_TG + lg);
  std::vector<int> order(chains.size());
  for (size_t i = 0; i < chains.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return chains[a].rows.size() > chains[b].rows.size(); });
  for (int ci : order) {
    const int len = (int)chains[ci].rows.size();
    int slot = -1;
    for (size_t s = 0; s < slots.size(); ++s)
      if (used[s] + len <= WAVE) {
        slot = (int)s;
        break;
      }
    if (slot < 0) {
      slots.emplace_back(WAVE, rd_make(R_NONE, 0, 0, 0));
      used.push_back(0);
      slot = (int)slots.size() - 1;
    }
    for (int j = 0; j < len; ++j) slots[slot][used[slot] + j] = chains[ci].rows[j];
    used[slot] += len;
  }
  if ((int)slots.size() > MAXNS) return fail(DEB_E_UNSUPPORTED, "state does not fit 6 lane slots");
  L.ns = (int)slots.size();
  L.rd.clear();
  for (auto &s : slots) L.rd.insert(L.rd.end(), s.begin(), s.end());
  return 0;
}

template <int NS>
int fill_params(Params<NS> &P, const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo,
                const double *tables, int64_t stride, const int32_t *desc, const double *q, const double *w, int nk,
                const double *kmodes, int nout, const double *tau_out) {
  std::memset(&P, 0, sizeof(P));
  P.cfg = *cfg;
  P.nq = cfg->nqmax;
  P.n = 12 + cfg->lmaxg + cfg->lmaxgp + cfg->lmaxr + cfg->nqmax * (cfg->lmaxnu + 1);
  P.n_cosmo = n_cosmo;
  P.nk = nk;
  P.nout = nout;
  P.table_stride = stride;
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    P.desc[s][0] = desc[2 * s];
    P.desc[s][1] = desc[2 * s + 1];
    if (desc[2 * s + 1] < 2 || desc[2 * s] < 0 || desc[2 * s] + 3 * (int64_t)desc[2 * s + 1] > stride)
      return fail(DEB_E_ARG, "spline descriptor outside the table stride");
  }
  for (int i = 0; i < P.nq; ++i) {
    P.q[i] = q[i];
    P.w[i] = w[i];
    P.dlnf[i] = -q[i] / (1.0 + std::exp(-q[i]));
  }
  P.cosmo = cosmo;
  P.tables = tables;
  P.kmodes = kmodes;
  P.tau_out = tau_out;
  for (int s = 0; s < NS; ++s)
    for (int l = 0; l < WAVE; ++l) P.rowdesc[s][l] = L.rd[(size_t)s * WAVE + l];
  return 0;
}

int check_common(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                 const int32_t *desc, const double *q, const double *w) {
  if (!cfg || !cosmo || !tables || !desc) return fail(DEB_E_ARG, "null argument");
  if (n_cosmo < 1) return fail(DEB_E_ARG, "n_cosmo must be >= 1");
  if (cfg->nqmax > 0 && (!q || !w)) return fail(DEB_E_ARG, "q/w required when nqmax > 0");
  if (cfg->solver != DEB_SOLVER_RODAS5) return fail(DEB_E_ARG, "unknown solver");
  if (!(cfg->rtol > 0) || !(cfg->atol >= 0)) return fail(DEB_E_ARG, "rtol must be > 0 and atol >= 0");
  if (cfg->max_steps < 1) return fail(DEB_E_ARG, "max_steps must be >= 1");
  return 0;
}

int hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) return fail(DEB_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

template <int NS>
int launch_forward(const Params<NS> &P, hipStream_t st) {
  const int64_t total = (int64_t)P.n_cosmo * P.nk;
  hipLaunchKernelGGL(deb_forward_kernel<NS>, dim3((unsigned)total), dim3(WAVE), 0, st, P);
  return hip_check(hipGetLastError(), "forward kernel launch");
}

template <int NS>
int run_forward(const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo,
                const double *tables, int64_t stride, const int32_t *desc, const double *q, const double *w, int nk,
                const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                int32_t *nsteps, void *ws, hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, n_cosmo, cosmo, tables, stride, desc, q, w, nk, kmodes, nout, tau_out))
    return r;
  P.y_out = y_out;
  P.status = status;
  P.nsteps = nsteps;
  (void)ws;
  return launch_forward<NS>(P, st);
}

template <int NS>
int run_hook(bool step, const deb_config_t *cfg, const Layout &L, const deb_cosmo_t *cosmo, const double *tables,
             int64_t stride, const int32_t *desc, const double *q, const double *w, int m, const int32_t *ci,
             const double *kk, const double *t0, const double *t1, const double *y0, double *y1, double *err,
             hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, 1, cosmo, tables, stride, desc, q, w, 1, nullptr, 1, nullptr)) return r;
  HookArgs<NS> H{ci, kk, t0, t1, y0, y1, err, m};
  if (step) hipLaunchKernelGGL(deb_step_kernel<NS>, dim3(m), dim3(WAVE), 0, st, P, H);
  else hipLaunchKernelGGL(deb_rhs_kernel<NS>, dim3(m), dim3(WAVE), 0, st, P, H);
  return hip_check(hipGetLastError(), step ? "step kernel launch" : "rhs kernel launch");
}

#define DEB_DISPATCH(NSV, CALL)                                      \
  switch (NSV) {                                                     \
    case 1: { constexpr int NS = 1; return CALL; }                   \
    case 2: { constexpr int NS = 2; return CALL; }                   \
    case 3: { constexpr int NS = 3; return CALL; }                   \
    case 4: { constexpr int NS = 4; return CALL; }                   \
    case 5: { constexpr int NS = 5; return CALL; }                   \
    case 6: { constexpr int NS = 6; return CALL; }                   \
    default: return fail(DEB_E_UNSUPPORTED, "unsupported slot count"); \
  }

}  // namespace deb

using namespace deb;

extern "C" {

int deb_abi_version(void) { return DEB_ABI_VERSION; }

const char *deb_last_error(void) { return g_err.c_str(); }

int deb_nvar(const deb_config_t *cfg) {
  if (!cfg) return fail(DEB_E_ARG, "null config");
  return 12 + cfg->lmaxg + cfg->lmaxgp + cfg->lmaxr + cfg->nqmax * (cfg->lmaxnu + 1);
}

int deb_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int nout, size_t *bytes) {
  if (!cfg || !bytes || n_cosmo < 1 || nk < 1 || nout < 1) return fail(DEB_E_ARG, "invalid argument");
  *bytes = 256;
  return 0;
}

int deb_solve_forward(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                      int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                      const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                      int32_t *nsteps, void *workspace, size_t workspace_bytes, void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !y_out || !status) return fail(DEB_E_ARG, "invalid mode arrays");
  if (!workspace || workspace_bytes < 256) return fail(DEB_E_ARG, "workspace smaller than deb_workspace_bytes");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_forward<NS>(cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes, nout,
                                      tau_out, y_out, status, nsteps, workspace, st)))
}

int deb_rhs(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
            const int32_t *desc, const double *q, const double *w, int m, const int32_t *ci, const double *kk,
            const double *tau, const double *y, double *f, void *stream) {
  if (int r = check_common(cfg, 1, cosmo, tables, desc, q, w)) return r;
  if (m < 1 || !ci || !kk || !tau || !y || !f) return fail(DEB_E_ARG, "invalid hook arrays");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_hook<NS>(false, cfg, L, cosmo, tables, table_stride, desc, q, w, m, ci, kk, tau, nullptr,
                                   y, f, nullptr, st)))
}

int deb_rodas5_step(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
                    const int32_t *desc, const double *q, const double *w, int m, const int32_t *ci,
                    const double *kk, const double *t0, const double *t1, const double *y0, double *y1, double *err,
                    void *stream) {
  if (int r = check_common(cfg, 1, cosmo, tables, desc, q, w)) return r;
  if (m < 1 || !ci || !kk || !t0 || !t1 || !y0 || !y1 || !err) return fail(DEB_E_ARG, "invalid hook arrays");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_hook<NS>(true, cfg, L, cosmo, tables, table_stride, desc, q, w, m, ci, kk, t0, t1, y0, y1,
                                   err, st)))
}

}  // extern "C"
