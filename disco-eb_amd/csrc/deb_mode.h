// Per-mode physics and the structured Rodas5 step (device side).
//
// Physics: model_synchronous (src/discoeb/perturbations.py:84-371), written as
// one row per lane: f_row = sub*y[lane-1] + diag*y[lane] + sup*y[lane+1]
//   + uh*h' + ueta*eta' + upi*Pi + leaf*y[partner]
// with the metric functionals h', eta' (perturbations.py:257-260), the
// polarization source Pi = F_2 + G_0 + G_2 (:286) and the a*h' accumulator
// f_1 (:252) obtained as wave reductions sum(c * y).
#pragma once

#include <type_traits>

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
  // input tangents (gradient kernels): dcos[ndir][24], dtab[ndir][stride]
  const double *dcos;
  const double *dtab;
  int64_t stride;
  int ndir;
};

// deb_cosmo_t field indices
enum : int {
  F_GRHOM = 0, F_GRHOG, F_GRHOR, F_OMEGAM, F_OMEGAB, F_OMEGADE, F_OMEGAK, F_NEFF, F_NMNU, F_AMNU, F_W0, F_WA,
  F_CS2DE, F_YHE, F_H0, F_TAUMIN
};
constexpr int COSMO_DOUBLES = 24;

// cosmology scalar f as a T-compatible parameter: plain double, or (for dn
// types) the value with its input tangents in components 1..ndir
template <class T>
__device__ __forceinline__ ParT<T> cv(const Ctx &cx, int f) {
  const double v = reinterpret_cast<const double *>(cx.c)[f];
  if constexpr (is_dn<T>::value) {
    T r;
    r = v;
#pragma unroll
    for (int j = 0; j + 1 < T::size; ++j)
      if (j < cx.ndir) r.d[j + 1] = cx.dcos[j * COSMO_DOUBLES + f];
    return r;
  } else {
    return v;
  }
}

// spline s at x: dn types also carry the knots' input tangents (x, y and S)
template <class T>
__device__ __forceinline__ T spl_at(Ctx &cx, int s, T xv) {
  if constexpr (is_dn<T>::value) {
    const Spl &sp = cx.sp[s];
    const int i = spl_interval(sp, vval(xv), cx.hint[s]);
    auto knot = [&](int part, int idx) {
      const double *base = part == 0 ? sp.x : (part == 1 ? sp.y : sp.S);
      T r;
      r = base[idx];
#pragma unroll
      for (int j = 0; j + 1 < T::size; ++j)
        if (j < cx.ndir) r.d[j + 1] = cx.dtab[j * cx.stride + sp.off + part * sp.n + idx];
      return r;
    };
    T x0 = knot(0, i), x1 = knot(0, i + 1);
    T h = x1 - x0;
    T t = (xv - x0) / h;
    T A = 1.0 - t;
    T B = t;
    return A * knot(1, i) + B * knot(1, i + 1) +
           ((A * A * A - A) * knot(2, i) + (B * B * B - B) * knot(2, i + 1)) * (h * h / 6.0);
  } else {
    return spl_eval(cx.sp[s], xv, cx.hint[s]);
  }
}

// rho_Q(a) / rho_Q(1) = a^{-3(1+w0+wa)} exp(3 (a-1) wa); for dn types the
// exponent's own tangent enters through exp((e - e0) log a)
template <class T>
__device__ __forceinline__ T rho_de(T a, T loga, ParT<T> w0, ParT<T> wa) {
  const double e0 = vval(-3.0 * (1.0 + w0 + wa));
  T r = dpow(a, e0) * dexp(3.0 * (a - 1.0) * wa);
  if constexpr (is_dn<T>::value) r = r * dexp((-3.0 * (1.0 + w0 + wa) - e0) * loga);
  return r;
}

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
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), OmegaDE = cv<T>(cx, F_OMEGADE);
  const P Omegak = cv<T>(cx, F_OMEGAK), Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU);
  const P w0 = cv<T>(cx, F_W0), wa = cv<T>(cx, F_WA), YHe = cv<T>(cx, F_YHE), H0 = cv<T>(cx, F_H0);
  Bg<T> b;
  b.a = a;
  T loga = dlog(a);
  T cs2a, xe, lrhonu;
  if constexpr (is_dn<T>::value) {
    cs2a = spl_at<T>(cx, DEB_SPL_CS2A_LOGA, loga);
    xe = spl_at<T>(cx, DEB_SPL_XE_LOGA, loga);
    lrhonu = spl_at<T>(cx, DEB_SPL_LOGRHONU_LOGA, loga);
  } else {
    const Spl sp3[3] = {cx.sp[DEB_SPL_CS2A_LOGA], cx.sp[DEB_SPL_XE_LOGA], cx.sp[DEB_SPL_LOGRHONU_LOGA]};
    int *const h3[3] = {&cx.hint[DEB_SPL_CS2A_LOGA], &cx.hint[DEB_SPL_XE_LOGA], &cx.hint[DEB_SPL_LOGRHONU_LOGA]};
    T r3[3];
    spl_eval3<T>(sp3, loga, h3, r3);
    cs2a = r3[0];
    xe = r3[1];
    lrhonu = r3[2];
  }
  b.cs2 = cs2a / a;
  T rhonu = dexp(lrhonu);
  b.pb43 = (4.0 / 3.0) * (grhog / (grhom * Omegab)) / a;
  b.w = w0 + wa * (1.0 - a);
  b.rhoQ = rho_de<T>(a, loga, w0, wa);
  T a2 = a * a;
  T grho = grhom * Omegam / a + (grhog + grhor * (Neff + Nmnu * rhonu)) / a2 + grhom * OmegaDE * b.rhoQ * a2 +
           grhom * Omegak;
  b.aH = dsqrt(grho / 3.0);
  T wprime = -wa * b.aH * a;
  b.ca2 = b.w - wprime / 3.0 / ((1.0 + b.w) + 1e-6) / b.aH;
  const P akthom = 2.3038921003709498e-9 * (1.0 - YHe) * Omegab * H0 * H0;
  b.opac = xe * akthom / a2;
  return b;
}

template <typename T>
__device__ __forceinline__ T nu_v(const Ctx &cx, T a, int qi) {
  T aq = a * (cv<T>(cx, F_AMNU) / cx.q[qi]);
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

// coefficient fields of a row
enum : int { C_SUB = 0, C_DIAG, C_SUP, C_UH, C_UETA, C_UPI, C_CH, C_CETA, C_CPI, C_C1, C_LEAF };

template <typename T, class Sink>
__device__ __forceinline__ void row_visit(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau, Sink sink) {
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), OmegaDE = cv<T>(cx, F_OMEGADE);
  const P Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU), cs2DE = cv<T>(cx, F_CS2DE);
  const double k = cx.k, k2 = k * k;
  const uint32_t t = rd_type(d);
  const double l = (double)rd_l(d);
  const int qi = (int)rd_q(d);
  const T a = b.a;
  const T a2 = a * a;
  const P Omegac = Omegam - Omegab;
  const T inv_aH = 1.0 / b.aH;
  // multipole couplings l/(2l+1) and (l+1)/(2l+1), and the massive-nu velocity
  // of the row's momentum bin: evaluated once here rather than in each of the
  // (divergent) row cases that use them
  const double il = 1.0 / (2.0 * l + 1.0);
  const double lsub = l * il, lsup = (l + 1.0) * il;
  const bool qrow = t >= R_Q0 && t <= R_QLMAX;
  T vq = tconst<T>(1.0);
  if (qrow) vq = nu_v<T>(cx, a, qi);
  switch (t) {
    case R_ETA:
      sink(std::integral_constant<int, C_UETA>{}, 1.0);
      sink(std::integral_constant<int, C_CH>{}, 2.0 * k2 * inv_aH);
      break;
    case R_DC:
      sink(std::integral_constant<int, C_SUP>{}, -1.0);
      sink(std::integral_constant<int, C_UH>{}, -0.5);
      sink(std::integral_constant<int, C_CH>{}, grhom * Omegac / a * inv_aH);
      sink(std::integral_constant<int, C_C1>{}, -a * (grhom * Omegac / a));
      break;
    case R_TC:
      sink(std::integral_constant<int, C_DIAG>{}, -b.aH);
      sink(std::integral_constant<int, C_CETA>{}, grhom * Omegac / a * (0.5 / k2));
      break;
    case R_DB:
      sink(std::integral_constant<int, C_SUP>{}, -1.0);
      sink(std::integral_constant<int, C_UH>{}, -0.5);
      sink(std::integral_constant<int, C_CH>{}, grhom * Omegab / a * inv_aH);
      sink(std::integral_constant<int, C_C1>{}, -a * (grhom * Omegab / a));
      break;
    case R_TB:
      sink(std::integral_constant<int, C_SUB>{}, k2 * b.cs2);
      sink(std::integral_constant<int, C_DIAG>{}, -b.aH - b.pb43 * b.opac);
      sink(std::integral_constant<int, C_SUP>{}, b.pb43 * b.opac);
      sink(std::integral_constant<int, C_CETA>{}, grhom * Omegab / a * (0.5 / k2));
      break;
    case R_DG:
      sink(std::integral_constant<int, C_LEAF>{}, -4.0 / 3.0);
      sink(std::integral_constant<int, C_UH>{}, -2.0 / 3.0);
      sink(std::integral_constant<int, C_CH>{}, grhog / a2 * inv_aH);
      sink(std::integral_constant<int, C_C1>{}, -a * (grhog / a2 + 3.0 * grhog / a2 / 3.0));
      break;
    case R_TG:
      sink(std::integral_constant<int, C_SUB>{}, b.opac);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -0.5 * k2);
      sink(std::integral_constant<int, C_LEAF>{}, 0.25 * k2);
      sink(std::integral_constant<int, C_CETA>{}, (4.0 / 3.0) * grhog / a2 * (0.5 / k2));
      break;
    case R_F2:
      sink(std::integral_constant<int, C_SUB>{}, 8.0 / 15.0);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -0.6 * k);
      sink(std::integral_constant<int, C_UH>{}, 4.0 / 15.0);
      sink(std::integral_constant<int, C_UETA>{}, 8.0 / 5.0);
      sink(std::integral_constant<int, C_UPI>{}, 0.1 * b.opac);
      sink(std::integral_constant<int, C_CPI>{}, 1.0);
      break;
    case R_FL:
      sink(std::integral_constant<int, C_SUB>{}, k * lsub);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -k * lsup);
      break;
    case R_FLMAX:
      sink(std::integral_constant<int, C_SUB>{}, k);
      sink(std::integral_constant<int, C_DIAG>{}, -(l + 1.0) * invtau - b.opac);
      break;
    case R_G0:
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -k);
      sink(std::integral_constant<int, C_UPI>{}, 0.5 * b.opac);
      sink(std::integral_constant<int, C_CPI>{}, 1.0);
      break;
    case R_GL:
      sink(std::integral_constant<int, C_SUB>{}, k * lsub);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -k * lsup);
      break;
    case R_G2:
      sink(std::integral_constant<int, C_SUB>{}, k * 2.0 / 5.0);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -k * 3.0 / 5.0);
      sink(std::integral_constant<int, C_UPI>{}, 0.1 * b.opac);
      sink(std::integral_constant<int, C_CPI>{}, 1.0);
      break;
    case R_GLMAX:
      sink(std::integral_constant<int, C_SUB>{}, k);
      sink(std::integral_constant<int, C_DIAG>{}, -(l + 1.0) * invtau - b.opac);
      break;
    case R_DR:
      sink(std::integral_constant<int, C_SUP>{}, -4.0 / 3.0);
      sink(std::integral_constant<int, C_UH>{}, -2.0 / 3.0);
      sink(std::integral_constant<int, C_CH>{}, grhor * Neff / a2 * inv_aH);
      sink(std::integral_constant<int, C_C1>{}, -a * (grhor * Neff / a2 + grhor * Neff / a2));
      break;
    case R_TR:
      sink(std::integral_constant<int, C_SUB>{}, 0.25 * k2);
      sink(std::integral_constant<int, C_SUP>{}, -0.5 * k2);
      sink(std::integral_constant<int, C_CETA>{}, (4.0 / 3.0) * Neff * grhor / a2 * (0.5 / k2));
      break;
    case R_FR2:
      sink(std::integral_constant<int, C_SUB>{}, 8.0 / 15.0);
      sink(std::integral_constant<int, C_SUP>{}, -0.6 * k);
      sink(std::integral_constant<int, C_UH>{}, 4.0 / 15.0);
      sink(std::integral_constant<int, C_UETA>{}, 8.0 / 5.0);
      break;
    case R_FRL:
      sink(std::integral_constant<int, C_SUB>{}, k * lsub);
      sink(std::integral_constant<int, C_SUP>{}, -k * lsup);
      break;
    case R_FRLMAX:
      sink(std::integral_constant<int, C_SUB>{}, k);
      sink(std::integral_constant<int, C_DIAG>{}, -(l + 1.0) * invtau);
      break;
    case R_Q0: {
      T kv = k * vq;
      sink(std::integral_constant<int, C_SUP>{}, -kv);
      sink(std::integral_constant<int, C_UH>{}, cx.dlnf[qi] / 6.0);
      sink(std::integral_constant<int, C_CH>{}, grhor * Nmnu * cx.w[qi] / vq / a2 * inv_aH);
      sink(std::integral_constant<int, C_C1>{}, -a * (grhor * Nmnu * cx.w[qi] / vq / a2 + grhor * Nmnu * cx.w[qi] * vq / a2));
      break;
    }
    case R_Q1: {
      T kv = k * vq;
      sink(std::integral_constant<int, C_SUB>{}, kv / 3.0);
      sink(std::integral_constant<int, C_SUP>{}, -2.0 * kv / 3.0);
      sink(std::integral_constant<int, C_CETA>{}, Nmnu * grhor * k * cx.w[qi] / a2 * (0.5 / k2));
      break;
    }
    case R_Q2: {
      T kv = k * vq;
      sink(std::integral_constant<int, C_SUB>{}, 2.0 * kv / 5.0);
      sink(std::integral_constant<int, C_SUP>{}, -3.0 * kv / 5.0);
      sink(std::integral_constant<int, C_UH>{}, -cx.dlnf[qi] / 15.0);
      sink(std::integral_constant<int, C_UETA>{}, -0.4 * cx.dlnf[qi]);
      break;
    }
    case R_QL: {
      T kv = k * vq;
      sink(std::integral_constant<int, C_SUB>{}, kv * lsub);
      sink(std::integral_constant<int, C_SUP>{}, -kv * lsup);
      break;
    }
    case R_QLMAX: {
      sink(std::integral_constant<int, C_SUB>{}, k * vq);
      sink(std::integral_constant<int, C_DIAG>{}, -(l + 1.0) * invtau);
      break;
    }
    case R_DQ:
      if (!cx.lambda) {
        const P cs2Q = cs2DE;
        T opw = 1.0 + b.w;
        sink(std::integral_constant<int, C_DIAG>{}, -3.0 * (cs2Q - b.w) * b.aH);
        sink(std::integral_constant<int, C_SUP>{}, -opw - 9.0 * opw * (cs2Q - b.ca2) * b.aH * b.aH / k2);
        sink(std::integral_constant<int, C_UH>{}, -0.5 * opw);
        T rq = grhom * OmegaDE * b.rhoQ * a2;
        sink(std::integral_constant<int, C_CH>{}, rq * inv_aH);
        sink(std::integral_constant<int, C_C1>{}, -a * (rq + 3.0 * cs2Q * rq));
      }
      break;
    case R_TQ:
      if (!cx.lambda) {
        const P cs2Q = cs2DE;
        T opw = 1.0 + b.w;
        sink(std::integral_constant<int, C_SUB>{}, cs2Q / opw * k2);
        sink(std::integral_constant<int, C_DIAG>{}, -(1.0 - 3.0 * cs2Q) * b.aH);
        T rq = grhom * OmegaDE * b.rhoQ * a2;
        sink(std::integral_constant<int, C_CETA>{}, opw * rq * (0.5 / k2));
        sink(std::integral_constant<int, C_C1>{}, -a * (3.0 * (cs2Q - b.ca2) * (3.0 * b.aH / k2) * opw * rq));
      }
      break;
    default:
      break;
  }
}


// all coefficients of a row (fields not set by the row are zero)
template <typename T>
__device__ __forceinline__ Coef<T> row_coef(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau) {
  T v[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) v[i] = 0.0;
  row_visit<T>(d, b, cx, invtau, [&](auto f, const auto &x) __attribute__((always_inline)) {
    v[decltype(f)::value] = v[decltype(f)::value] + x;
  });
  Coef<T> r;
  r.sub = v[C_SUB]; r.diag = v[C_DIAG]; r.sup = v[C_SUP];
  r.uh = v[C_UH]; r.ueta = v[C_UETA]; r.upi = v[C_UPI];
  r.ch = v[C_CH]; r.ceta = v[C_CETA]; r.cpi = v[C_CPI]; r.c1 = v[C_C1]; r.leaf = v[C_LEAF];
  return r;
}

}  // namespace deb
