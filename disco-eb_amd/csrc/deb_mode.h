// Per-mode physics and the structured Rodas5 step (device side).
//
// Physics: model_synchronous (src/discoeb/perturbations.py:84-371), written as
// one row per lane: f_row = sub*y[lane-1] + diag*y[lane] + sup*y[lane+1]
//   + uh*h' + ueta*eta' + upi*Pi + leaf*y[partner]
// with the metric functionals h', eta' (perturbations.py:257-260), the
// polarization source Pi = F_2 + G_0 + G_2 (:286) and the a*h' accumulator
// f_1 (:252) obtained as wave reductions sum(c * y).
#pragma once

#include "deb_device.h"

namespace deb {

struct Ctx {
  const deb_cosmo_t *c;
  Spl sp[DEB_NSPLINE];
  double k;
  int nq;
  const double *q, *w, *dlnf;
  int hint[DEB_NSPLINE];
  bool lambda;  // w0 == -1 and wa == 0: dark-energy perturbations frozen
};

template <typename T>
struct Bg {
  T a, aH, opac, cs2, pb43, rhoQ, w, ca2;
};

// massive-nu velocity v = 1/sqrt(1 + (a m_nu / q)^2) of momentum bin qi
template <typename T>
__device__ __forceinline__ T nu_v(const Ctx &cx, T a, int qi);



// background at scale factor a (get_aprimeoa background.py:100-122 and the
// thermodynamics lookups of perturbations.py:176-218)
template <typename T>
__device__ __forceinline__ Bg<T> background(Ctx &cx, T a) {
  const deb_cosmo_t &c = *cx.c;
  Bg<T> b;
  b.a = a;
  T loga = dlog(a);
  b.cs2 = spl_eval(cx.sp[DEB_SPL_CS2A_LOGA], loga, cx.hint[DEB_SPL_CS2A_LOGA]) / a;
  T xe = spl_eval(cx.sp[DEB_SPL_XE_LOGA], loga, cx.hint[DEB_SPL_XE_LOGA]);
  T rhonu = dexp(spl_eval(cx.sp[DEB_SPL_LOGRHONU_LOGA], loga, cx.hint[DEB_SPL_LOGRHONU_LOGA]));
  b.pb43 = (4.0 / 3.0) * (c.grhog / (c.grhom * c.Omegab)) / a;
  b.w = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  b.rhoQ = dpow(a, -3.0 * (1.0 + c.w_DE_0 + c.w_DE_a)) * dexp(3.0 * (a - 1.0) * c.w_DE_a);
  T a2 = a * a;
  T grho = c.grhom * c.Omegam / a + (c.grhog + c.grhor * (c.Neff + c.Nmnu * rhonu)) / a2 +
           c.grhom * c.OmegaDE * b.rhoQ * a2 + c.grhom * c.Omegak;
  b.aH = dsqrt(grho / 3.0);
  T wprime = -c.w_DE_a * b.aH * a;
  b.ca2 = b.w - wprime / 3.0 / ((1.0 + b.w) + 1e-6) / b.aH;
  double akthom = 2.3038921003709498e-9 * (1.0 - c.YHe) * c.Omegab * c.H0 * c.H0;
  b.opac = xe * akthom / a2;
  return b;
}

template <typename T>
__device__ __forceinline__ T nu_v(const Ctx &cx, T a, int qi) {
  T aq = a * (cx.c->amnu / cx.q[qi]);
  return 1.0 / dsqrt(1.0 + aq * aq);
}

template <typename T>
struct Coef {
  T sub, diag, sup;   // local tridiagonal part of J
  T uh, ueta, upi;    // where h', eta', Pi enter the row
  T ch, ceta, cpi;    // contributions of this variable to h', eta', Pi
  T c1;               // contribution to f_1 = -(dgrho + 3 dgpres) a
  T leaf;             // coupling to the leaf partner (delta_gamma <-> theta_gamma)
};

template <typename T>
__device__ __forceinline__ Coef<T> row_coef(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau) {
  const deb_cosmo_t &c = *cx.c;
  const double k = cx.k, k2 = k * k;
  Coef<T> r;
  r.sub = 0.0; r.diag = 0.0; r.sup = 0.0;
  r.uh = 0.0; r.ueta = 0.0; r.upi = 0.0;
  r.ch = 0.0; r.ceta = 0.0; r.cpi = 0.0; r.c1 = 0.0; r.leaf = 0.0;
  const uint32_t t = rd_type(d);
  const double l = (double)rd_l(d);
  const int qi = (int)rd_q(d);
  const T a = b.a;
  const T a2 = a * a;
  const double Omegac = c.Omegam - c.Omegab;
  const T inv_aH = 1.0 / b.aH;
  switch (t) {
    case R_ETA:
      r.ueta = 1.0;
      r.ch = 2.0 * k2 * inv_aH;
      break;
    case R_DC:
      r.sup = -1.0;
      r.uh = -0.5;
      r.ch = c.grhom * Omegac / a * inv_aH;
      r.c1 = -a * (c.grhom * Omegac / a);
      break;
    case R_TC:
      r.diag = -b.aH;
      r.ceta = c.grhom * Omegac / a * (0.5 / k2);
      break;
    case R_DB:
      r.sup = -1.0;
      r.uh = -0.5;
      r.ch = c.grhom * c.Omegab / a * inv_aH;
      r.c1 = -a * (c.grhom * c.Omegab / a);
      break;
    case R_TB:
      r.sub = k2 * b.cs2;
      r.diag = -b.aH - b.pb43 * b.opac;
      r.sup = b.pb43 * b.opac;
      r.ceta = c.grhom * c.Omegab / a * (0.5 / k2);
      break;
    case R_DG:
      r.leaf = -4.0 / 3.0;
      r.uh = -2.0 / 3.0;
      r.ch = c.grhog / a2 * inv_aH;
      r.c1 = -a * (c.grhog / a2 + 3.0 * c.grhog / a2 / 3.0);
      break;
    case R_TG:
      r.sub = b.opac;
      r.diag = -b.opac;
      r.sup = -0.5 * k2;
      r.leaf = 0.25 * k2;
      r.ceta = (4.0 / 3.0) * c.grhog / a2 * (0.5 / k2);
      break;
    case R_F2:
      r.sub = 8.0 / 15.0;
      r.diag = -b.opac;
      r.sup = -0.6 * k;
      r.uh = 4.0 / 15.0;
      r.ueta = 8.0 / 5.0;
      r.upi = 0.1 * b.opac;
      r.cpi = 1.0;
      break;
    case R_FL:
      r.sub = k * l / (2.0 * l + 1.0);
      r.diag = -b.opac;
      r.sup = -k * (l + 1.0) / (2.0 * l + 1.0);
      break;
    case R_FLMAX:
      r.sub = k;
      r.diag = -(l + 1.0) * invtau - b.opac;
      break;
    case R_G0:
      r.diag = -b.opac;
      r.sup = -k;
      r.upi = 0.5 * b.opac;
      r.cpi = 1.0;
      break;
    case R_GL:
      r.sub = k * l / (2.0 * l + 1.0);
      r.diag = -b.opac;
      r.sup = -k * (l + 1.0) / (2.0 * l + 1.0);
      break;
    case R_G2:
      r.sub = k * 2.0 / 5.0;
      r.diag = -b.opac;
      r.sup = -k * 3.0 / 5.0;
      r.upi = 0.1 * b.opac;
      r.cpi = 1.0;
      break;
    case R_GLMAX:
      r.sub = k;
      r.diag = -(l + 1.0) * invtau - b.opac;
      break;
    case R_DR:
      r.sup = -4.0 / 3.0;
      r.uh = -2.0 / 3.0;
      r.ch = c.grhor * c.Neff / a2 * inv_aH;
      r.c1 = -a * (c.grhor * c.Neff / a2 + c.grhor * c.Neff / a2);
      break;
    case R_TR:
      r.sub = 0.25 * k2;
      r.sup = -0.5 * k2;
      r.ceta = (4.0 / 3.0) * c.Neff * c.grhor / a2 * (0.5 / k2);
      break;
    case R_FR2:
      r.sub = 8.0 / 15.0;
      r.sup = -0.6 * k;
      r.uh = 4.0 / 15.0;
      r.ueta = 8.0 / 5.0;
      break;
    case R_FRL:
      r.sub = k * l / (2.0 * l + 1.0);
      r.sup = -k * (l + 1.0) / (2.0 * l + 1.0);
      break;
    case R_FRLMAX:
      r.sub = k;
      r.diag = -(l + 1.0) * invtau;
      break;
    case R_Q0: {
      T vq = nu_v<T>(cx, a, qi);
      T kv = k * vq;
      r.sup = -kv;
      r.uh = cx.dlnf[qi] / 6.0;
      r.ch = c.grhor * c.Nmnu * cx.w[qi] / vq / a2 * inv_aH;
      r.c1 = -a * (c.grhor * c.Nmnu * cx.w[qi] / vq / a2 + c.grhor * c.Nmnu * cx.w[qi] * vq / a2);
      break;
    }
    case R_Q1: {
      T vq = nu_v<T>(cx, a, qi);
      T kv = k * vq;
      r.sub = kv / 3.0;
      r.sup = -2.0 * kv / 3.0;
      r.ceta = c.Nmnu * c.grhor * k * cx.w[qi] / a2 * (0.5 / k2);
      break;
    }
    case R_Q2: {
      T vq = nu_v<T>(cx, a, qi);
      T kv = k * vq;
      r.sub = 2.0 * kv / 5.0;
      r.sup = -3.0 * kv / 5.0;
      r.uh = -cx.dlnf[qi] / 15.0;
      r.ueta = -0.4 * cx.dlnf[qi];
      break;
    }
    case R_QL: {
      T vq = nu_v<T>(cx, a, qi);
      T kv = k * vq;
      r.sub = kv * (l / (2.0 * l + 1.0));
      r.sup = -kv * ((l + 1.0) / (2.0 * l + 1.0));
      break;
    }
    case R_QLMAX: {
      T vq = nu_v<T>(cx, a, qi);
      r.sub = k * vq;
      r.diag = -(l + 1.0) * invtau;
      break;
    }
    case R_DQ:
      if (!cx.lambda) {
        const double cs2Q = c.cs2_DE;
        T opw = 1.0 + b.w;
        r.diag = -3.0 * (cs2Q - b.w) * b.aH;
        r.sup = -opw - 9.0 * opw * (cs2Q - b.ca2) * b.aH * b.aH / k2;
        r.uh = -0.5 * opw;
        T rq = c.grhom * c.OmegaDE * b.rhoQ * a2;
        r.ch = rq * inv_aH;
        r.c1 = -a * (rq + 3.0 * cs2Q * rq);
      }
      break;
    case R_TQ:
      if (!cx.lambda) {
        const double cs2Q = c.cs2_DE;
        T opw = 1.0 + b.w;
        r.sub = cs2Q / opw * k2;
        r.diag = -(1.0 - 3.0 * cs2Q) * b.aH;
        T rq = c.grhom * c.OmegaDE * b.rhoQ * a2;
        r.ceta = opw * rq * (0.5 / k2);
        r.c1 = -a * (3.0 * (cs2Q - b.ca2) * (3.0 * b.aH / k2) * opw * rq);
      }
      break;
    default:
      break;
  }
  return r;
}

}  // namespace deb
