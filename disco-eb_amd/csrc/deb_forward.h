// Device side of the forward solve: start time, initial conditions, output
// fields and the adaptive Rodas5 loop of one (cosmology, k) item.
#pragma once

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

// convert_to_output_variables (perturbations.py:374-523).  Y(i) returns the
// state entry of reference index i as a T (staged in LDS by the caller);
// sink(f, value) receives output field f.
template <class T, class Yf, class Sink>
__device__ __forceinline__ void output_fields_t(Ctx &cx, Yf Y, int lg, int lgp, int lr, int nq, int n, Sink sink) {
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), OmegaDE = cv<T>(cx, F_OMEGADE);
  const P Omegak = cv<T>(cx, F_OMEGAK), Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU);
  const P amnu = cv<T>(cx, F_AMNU), w0 = cv<T>(cx, F_W0), wa = cv<T>(cx, F_WA);
  const double k = cx.k, k2 = k * k;
  const P Omegac = Omegam - Omegab;
  const int iq0 = 10 + lg + lgp + lr, iq1 = iq0 + nq;
  T a = Y(0), eta = Y(2), deltac = Y(3), thetac = Y(4), deltab = Y(5), thetab = Y(6);
  T deltag = Y(7), thetag = Y(8);
  T deltar = Y(9 + lg + lgp), thetar = Y(10 + lg + lgp);
  T la = dlog(a);
  T rhonu = dexp(spl_at<T>(cx, DEB_SPL_LOGRHONU_LOGA, la));
  T pnu = dexp(spl_at<T>(cx, DEB_SPL_LOGPNU_LOGA, la));
  T drhonu = tconst<T>(0.0), fnu = tconst<T>(0.0);
  for (int i = 0; i < nq; ++i) {
    T aq = a * amnu / cx.q[i];
    T v = 1.0 / dsqrt(1.0 + aq * aq);
    drhonu = drhonu + cx.w[i] * Y(iq0 + i) / v;
    fnu = fnu + cx.w[i] * Y(iq1 + i);
  }
  T deltanu = drhonu / rhonu;
  T thetanu = k * fnu / (rhonu + pnu);
  T deltaq = Y(n - 2), thetaq = Y(n - 1);
  T wQ = w0 + wa * (1.0 - a);
  T rhoQ = rho_de<T>(a, la, w0, wa);
  T a2 = a * a;
  T rppQ = (1.0 + wQ) * rhoQ * grhom * OmegaDE * thetaq * a2;
  T grho = grhom * Omegam / a + (grhog + grhor * (Neff + Nmnu * rhonu)) / a2 + grhom * OmegaDE * rhoQ * a2 +
           grhom * Omegak;
  T aH = dsqrt(grho / 3.0);
  T dgrho = grhom * (Omegac * deltac + Omegab * deltab) / a + (grhog * deltag + grhor * (Neff * deltar + Nmnu * drhonu)) / a2 +
            grhom * OmegaDE * deltaq * rhoQ * a2;
  T dgtheta = grhom * (Omegac * thetac + Omegab * thetab) / a +
              4.0 / 3.0 * (grhog * thetag + Neff * grhor * thetar) / a2 + Nmnu * grhor * k * fnu / a2 + rppQ;
  T hprime = (2.0 * k2 * eta + dgrho) / aH;
  T etaprime = 0.5 * dgtheta / k2;
  T alpha = (hprime + 6.0 * etaprime) / 2.0 / k2;
  T mden = grhom * Omegam / a + grhor * Nmnu * rhonu / a2;
  T deltam = (grhom * (Omegac * deltac + Omegab * deltab) / a + grhor * Nmnu * drhonu / a2) / mden;
  T thetam = (grhom * (Omegac * thetac + Omegab * thetab) / a + Nmnu * grhor * k * fnu / a2) / (3.0 * mden);
  T deltabc = (grhom * (Omegac * deltac + Omegab * deltab) / a) / (grhom * Omegam / a);
  T thetabc = (grhom * (Omegac * thetac + Omegab * thetab) / a) / (3.0 * (grhom * Omegam / a) / a2);
  thetam = thetam + alpha * k2;
  thetabc = thetabc + alpha * k2;
  sink(0, eta); sink(1, etaprime); sink(2, hprime); sink(3, alpha);
  sink(4, deltam); sink(5, thetam / aH); sink(6, deltabc); sink(7, thetabc / aH);
  sink(8, deltac); sink(9, thetac / aH); sink(10, deltab); sink(11, thetab / aH);
  sink(12, deltag); sink(13, thetag / aH); sink(14, deltar); sink(15, thetar / aH);
  sink(16, deltanu); sink(17, thetanu / aH); sink(18, deltaq); sink(19, thetaq / aH);
}

template <int NS>
__device__ __forceinline__ void make_ctx(Ctx &cx, const Params<NS> &P, int ci, double k) {
  cx.c = P.cosmo + ci;
  const double *tb = P.tables + (int64_t)ci * P.table_stride;
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    int off = P.desc[s][0], n = P.desc[s][1];
    cx.sp[s] = {tb + off, tb + off + n, tb + off + 2 * n, n, off};
    cx.hint[s] = -1;
  }
  cx.dcos = nullptr;
  cx.dtab = nullptr;
  cx.stride = P.table_stride;
  cx.ndir = 0;
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
    output_fields_t<double>(cx, [&](int i) { return lds[i]; }, P.cfg.lmaxg, P.cfg.lmaxgp, P.cfg.lmaxr, P.nq, P.n,
                            [&](int f, double v) { o[f] = v; });
#pragma unroll
    for (int j = 0; j < DEB_NOUT_FIELDS; ++j)
      if (lane == j) dst[j] = o[j];
  }
  __builtin_amdgcn_s_barrier();
}

__device__ __forceinline__ unsigned long long readfirst_u64(unsigned long long x) {
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(x & 0xffffffffull));
  unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// records the accepted states of one mode on the checkpoint tape: every
// state's time, and the state itself at the start of each segment of C steps
template <int NS>
struct TapeWriter {
  TapeView tv;
  long long mode;
  long long cur;
  int cnt;
  bool full;
  __device__ __forceinline__ double *times(long long chunk) const {
    return reinterpret_cast<double *>(tv.pool + chunk * tv.chunk_bytes + 64);
  }
  __device__ __forceinline__ void put(double t, const double (&y)[NS]) {
    if (full) return;
    const int lane = lane_id();
    const int j = cnt % tv.ckpt;
    if (j == 0) {
      // state cnt ends the previous segment and starts a new one
      if (cnt > 0 && lane == 0) times(cur)[tv.ckpt] = t;
      unsigned long long c = 0;
      if (lane == 0) c = atomicAdd(tv.counter, 1ull);
      c = readfirst_u64(c);
      if ((long long)c >= tv.capacity) {
        full = true;
        return;
      }
      char *chunk = tv.pool + (long long)c * tv.chunk_bytes;
      if (lane == 0) *reinterpret_cast<long long *>(chunk) = cur;
      cur = (long long)c;
      double *yy = reinterpret_cast<double *>(chunk + tv.state_off);
#pragma unroll
      for (int s = 0; s < NS; ++s) yy[s * WAVE + lane] = y[s];
    }
    if (lane == 0) times(cur)[j] = t;
    ++cnt;
  }
  __device__ __forceinline__ void finish() {
    if (lane_id() == 0) {
      tv.last[mode] = cur;
      tv.count[mode] = full ? -1 : cnt;
    }
  }
};

// diffrax replaces a NaN error by inf: a non-finite norm rejects the step and
// shrinks it by factormin
__device__ __forceinline__ double norm_or_inf(double nrm) { return nrm <= 1.7976931348623157e308 ? nrm : INFINITY; }

// true on every lane if any lane holds a non-finite state entry
template <int NS>
__device__ __forceinline__ bool any_nonfinite(const double (&y)[NS]) {
  bool bad = false;
#pragma unroll
  for (int s = 0; s < NS; ++s) bad = bad || !isfinite(y[s]);
  return __ballot(bad) != 0;
}

// outputs a failed mode never reached hold inf (diffrax fills unreached
// SaveAt entries with inf)
__device__ __forceinline__ void fill_unreached(double *dst, int iout, int nout, int width) {
  for (int i = iout * width + lane_id(); i < nout * width; i += WAVE) dst[i] = INFINITY;
}

__device__ __forceinline__ int mode_levels(const deb_config_t &cf, int nq) {
  int maxlen = cf.lmaxg + 2;
  maxlen = max(maxlen, cf.lmaxgp + 1);
  maxlen = max(maxlen, cf.lmaxr + 1);
  if (nq > 0) maxlen = max(maxlen, cf.lmaxnu + 1);
  int nl = 0;
  while ((1 << nl) < maxlen) ++nl;
  return nl;
}

// evolve_one_mode (perturbations.py:728-781) for one (cosmology, k) item;
// with TAPE every accepted state is also written to the checkpoint tape
template <int NS, bool TAPE = false>
__device__ __forceinline__ void integrate_mode(const Params<NS> &P, int ci, int ki, double *lds, lds_double *pcr,
                                               const TapeView *tape = nullptr) {
  const int lane = lane_id();
  Ctx cx;
  const double k = P.kmodes[(int64_t)ci * P.nk + ki];
  make_ctx<NS>(cx, P, ci, k);
  __shared__ ModeU mu;  // the wave-uniform step factors
  Mode<NS> m;
  m.pcr = pcr;
  m.u = (lds_mode_u *)&mu;
#pragma unroll
  for (int s = 0; s < NS; ++s) m.rd[s] = P.rowdesc[s][lane];
  m.nlev = mode_levels(P.cfg, P.nq);
  // stage vectors in the kernel's LDS (lds holds fwd_lds_doubles<NS>())
  lds_double *const kl = stage_in_lds<NS>() ? (lds_double *)lds : nullptr;
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
  TapeWriter<NS> tw;
  if constexpr (TAPE) {
    tw.tv = *tape;
    tw.mode = (long long)ci * P.nk + ki;
    tw.cur = -1;
    tw.cnt = 0;
    tw.full = false;
    tw.put(t0, y);
  }
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
    rodas5_step<NS>(m, cx, tprev, dt, y, y1, err, kl);
    // filtered, weighted RMS of the scaled error: lanes outside the filter
    // never enter it, so a non-finite unfiltered entry cannot poison it
    const double sc = err[0] / (cf.atol + fmax(fabs(y[0]), fabs(y1[0])) * cf.rtol);
    const double x = ew != 0.0 ? sc * ew : 0.0;
    const double nrm = norm_or_inf(sqrt(wave_sum1(x * x) / 6.0));
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
      if constexpr (TAPE) tw.put(tnext, y);
    }
    double nt0 = keep ? tnext : tprev;
    double nt1 = nt0 + dtn;
    if (nt1 > t1 - 1e-10) nt1 = keep ? t1 : nt0 + 0.5 * (t1 - nt0);
    tprev = nt0;
    tnext = nt1;
    ++steps;
    if (keep ? any_nonfinite<NS>(y) : !(nt1 > nt0)) {
      // a non-finite accepted state, or a step size shrunk to nothing
      status = DEB_STATUS_NONFINITE;
      break;
    }
  }
  if (status != DEB_STATUS_OK) fill_unreached(ydst, iout, P.nout, width);
  if constexpr (TAPE) {
    tw.finish();
    if (status == DEB_STATUS_OK && tw.full) status = DEB_STATUS_TAPE_FULL;
  }
  if (lane == 0) {
    P.status[(int64_t)ci * P.nk + ki] = status;
    if (P.nsteps) P.nsteps[(int64_t)ci * P.nk + ki] = steps;
  }
}

// evolve_modes_batched (perturbations.py:820-922) for one group of B
// consecutive k (kb .. kb + B - 1) of cosmology ci.  One step sequence serves
// the group: its start time is the minimum of the modes' start times
// (determine_start_times_per_batch :786-803) and the controller sees the RMS
// over all modes' filtered, weighted errors (rms_norm_filtered_batched
// :690-694).  The wavefront steps the modes one after the other; each mode's
// state and proposal live in ws (2 * NS * WAVE doubles per mode, lane-owned).
template <int NS>
__device__ __forceinline__ void integrate_batch(const Params<NS> &P, int ci, int kb, int B, double *lds,
                                                lds_double *pcr, double *ws) {
  const int lane = lane_id();
  const int64_t mode0 = (int64_t)ci * P.nk + kb;
  Ctx cx;
  make_ctx<NS>(cx, P, ci, P.kmodes[mode0]);
  __shared__ ModeU mu;  // the wave-uniform step factors
  Mode<NS> m;
  m.pcr = pcr;
  m.u = (lds_mode_u *)&mu;
#pragma unroll
  for (int s = 0; s < NS; ++s) m.rd[s] = P.rowdesc[s][lane];
  m.nlev = mode_levels(P.cfg, P.nq);
  // stage vectors in the kernel's LDS (lds holds fwd_lds_doubles<NS>())
  lds_double *const kl = stage_in_lds<NS>() ? (lds_double *)lds : nullptr;
  auto state = [&](int b, int which) -> double * {
    return ws + ((mode0 + b) * 2 + which) * (int64_t)(NS * WAVE) + lane;
  };
  const double *tout = P.tau_out + (int64_t)ci * P.nout;
  double tmin = tout[0], tmax = tout[0];
  for (int i = 1; i < P.nout; ++i) {
    tmin = fmin(tmin, tout[i]);
    tmax = fmax(tmax, tout[i]);
  }
  double t0 = start_time(cx, tmin);
  for (int b = 1; b < B; ++b) {
    cx.k = P.kmodes[mode0 + b];
    t0 = fmin(t0, start_time(cx, tmin));
  }
  const double t1 = tmax;
  for (int b = 0; b < B; ++b) {
    cx.k = P.kmodes[mode0 + b];
    double y[NS];
    initial_conditions<NS>(cx, m.rd, t0, y);
    double *dst = state(b, 0);
#pragma unroll
    for (int s = 0; s < NS; ++s) dst[s * WAVE] = y[s];
  }
  const deb_config_t &cf = P.cfg;
  const double eo = 5.0;
  const double coeff1 = (cf.icoeff + cf.pcoeff + cf.dcoeff) / eo;
  const double coeff2 = -(cf.pcoeff + 2.0 * cf.dcoeff) / eo;
  const double coeff3 = cf.dcoeff / eo;
  double tprev = t0;
  double tnext = fmin(t0 + fmin(t0 / 4.0, 0.5 * (t1 - t0)), t1);
  double prev_inv = 1.0, prev_prev_inv = 1.0;
  int iout = 0, steps = 0;
  int status = DEB_STATUS_OK;
  const int width = cf.return_full ? P.n : DEB_NOUT_FIELDS;
  while (tprev < t1) {
    if (steps >= cf.max_steps) {
      status = DEB_STATUS_MAX_STEPS;
      break;
    }
    const double dt = tnext - tprev;
    double e2 = 0.0;
    for (int b = 0; b < B; ++b) {
      const double k = P.kmodes[mode0 + b];
      cx.k = k;
      double y[NS], y1[NS], err[NS];
      const double *src = state(b, 0);
#pragma unroll
      for (int s = 0; s < NS; ++s) y[s] = src[s * WAVE];
      rodas5_step<NS>(m, cx, tprev, dt, y, y1, err, kl);
      // error-norm weights: indices (a, eta, delta_c, delta_b, theta_b, delta_gamma) of slot 0
      double ew = 0.0;
      if (lane == L_A || lane == L_DC || lane == L_DB || lane == L_DG) ew = 1.0;
      if (lane == L_ETA) ew = k * k;
      if (lane == L_TB) ew = 1.0 / (k * k);
      const double sc = err[0] / (cf.atol + fmax(fabs(y[0]), fabs(y1[0])) * cf.rtol);
      const double x = ew != 0.0 ? sc * ew : 0.0;
      e2 += wave_sum1(x * x);
      double *dst = state(b, 1);
#pragma unroll
      for (int s = 0; s < NS; ++s) dst[s * WAVE] = y1[s];
    }
    const double nrm = norm_or_inf(sqrt(e2 / (6.0 * B)));
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
        const double th = (tout[iout] - tprev) / dt;
        for (int b = 0; b < B; ++b) {
          cx.k = P.kmodes[mode0 + b];
          const double *y0p = state(b, 0), *y1p = state(b, 1);
          double ys[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) ys[s] = y0p[s * WAVE] + th * (y1p[s * WAVE] - y0p[s * WAVE]);
          write_outputs<NS>(P, cx, m.rd, ys, lds, P.y_out + ((mode0 + b) * P.nout + iout) * (int64_t)width);
        }
        ++iout;
      }
      prev_prev_inv = prev_inv;
      prev_inv = inv;
      bool bad = false;
      for (int b = 0; b < B; ++b) {
        const double *src = state(b, 1);
        double *dst = state(b, 0);
        double y[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          y[s] = src[s * WAVE];
          dst[s * WAVE] = y[s];
        }
        bad = bad || any_nonfinite<NS>(y);
      }
      if (bad) status = DEB_STATUS_NONFINITE;
    }
    double nt0 = keep ? tnext : tprev;
    double nt1 = nt0 + dtn;
    if (nt1 > t1 - 1e-10) nt1 = keep ? t1 : nt0 + 0.5 * (t1 - nt0);
    tprev = nt0;
    tnext = nt1;
    ++steps;
    if (!keep && !(nt1 > nt0)) status = DEB_STATUS_NONFINITE;
    if (status != DEB_STATUS_OK) break;
  }
  if (status != DEB_STATUS_OK)
    for (int b = 0; b < B; ++b)
      fill_unreached(P.y_out + (mode0 + b) * P.nout * (int64_t)width, iout, P.nout, width);
  if (lane == 0) {
    for (int b = 0; b < B; ++b) {
      P.status[mode0 + b] = status;
      if (P.nsteps) P.nsteps[mode0 + b] = steps;
    }
  }
}

// One workgroup (one wavefront) per (cosmology, k) item; the hardware
// dispatcher hands the next item to whichever SIMD frees up first, and items
// are numbered so that the largest k (most steps) start first.
template <int NS>
__global__ void __launch_bounds__(WAVE) deb_forward_kernel(const Params<NS> P) {
  __shared__ double lds[fwd_lds_doubles<NS>()];
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * P.nk) return;
  const int ki = P.nk - 1 - item / P.n_cosmo;
  const int ci = item % P.n_cosmo;
  integrate_mode<NS>(P, ci, ki, lds, (lds_double *)pcr);
}

}  // namespace deb
