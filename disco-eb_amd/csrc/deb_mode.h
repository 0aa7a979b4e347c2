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
// (one exponential of -3 (1 + w0 + wa) log a + 3 (a - 1) wa; the exponent
// carries the parameters' tangents for dn types)
template <class T>
__device__ __forceinline__ T rho_de(T a, T loga, ParT<T> w0, ParT<T> wa) {
  return dexp((-3.0 * (1.0 + w0 + wa)) * loga + 3.0 * (a - 1.0) * wa);
}

// 1/sqrt(x): refined hardware estimate for plain doubles
template <class T>
__device__ __forceinline__ T rsqrt_t(T x) {
  if constexpr (std::is_same<T, double>::value) return rsq_nr(x);
  else return 1.0 / dsqrt(x);
}

// 1/x: refined hardware estimate for plain doubles
template <class T>
__device__ __forceinline__ T rcp_t(T x) {
  if constexpr (std::is_same<T, double>::value) return rcp_nr(x);
  else return 1.0 / x;
}

// background quantities at a, with the reciprocals 1/a, 1/a^2 and 1/(aH) that
// the row coefficients use (taken once per evaluation, not per row)
template <typename T>
struct Bg {
  T a, aH, opac, cs2, pb43, rhoQ, w, ca2;
  T ia, ia2, iaH;
};



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
  b.ia = 1.0 / a;
  b.ia2 = b.ia * b.ia;
  b.cs2 = cs2a * b.ia;
  T rhonu = dexp(lrhonu);
  b.pb43 = (4.0 / 3.0) * (grhog / (grhom * Omegab)) * b.ia;
  b.w = w0 + wa * (1.0 - a);
  b.rhoQ = rho_de<T>(a, loga, w0, wa);
  T a2 = a * a;
  T grho = grhom * Omegam * b.ia + (grhog + grhor * (Neff + Nmnu * rhonu)) * b.ia2 + grhom * OmegaDE * b.rhoQ * a2 +
           grhom * Omegak;
  b.aH = dsqrt(grho * (1.0 / 3.0));
  b.iaH = 1.0 / b.aH;
  T wprime = -wa * b.aH * a;
  b.ca2 = b.w - wprime * (1.0 / 3.0) / ((1.0 + b.w) + 1e-6) * b.iaH;
  const P akthom = 2.3038921003709498e-9 * (1.0 - YHe) * Omegab * H0 * H0;
  b.opac = xe * akthom * b.ia2;
  return b;
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

// The coefficient fields of one row (model_synchronous, perturbations.py:
// 84-371), each passed to sink(field, value).  Every case sets each of its
// fields once.  The formulas use the evaluation's reciprocals (b.ia, b.ia2,
// b.iaH) and per-mode constants, so a case body holds multiplications only:
// the divergent switch runs every case present in a slot, and a division per
// case multiplied its cost.
template <typename T, class Sink>
__device__ __forceinline__ void row_visit(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau, Sink sink) {
  using P = ParT<T>;
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB), OmegaDE = cv<T>(cx, F_OMEGADE);
  const P Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU), cs2DE = cv<T>(cx, F_CS2DE);
  const double k = cx.k, k2 = k * k;
  const double ik2 = 1.0 / k2, hk2 = 0.5 * ik2;  // wave-uniform
  const uint32_t t = rd_type(d);
  const double l = (double)rd_l(d);
  const int qi = (int)rd_q(d);
  const T a = b.a;
  const T ia = b.ia, ia2 = b.ia2, iaH = b.iaH;
  const P Omegac = Omegam - Omegab;
  // multipole couplings l/(2l+1) and (l+1)/(2l+1), and the massive-nu velocity
  // of the row's momentum bin: evaluated once here rather than in each of the
  // (divergent) row cases that use them
  const double il = rcp_nr(2.0 * l + 1.0);
  const double lsub = l * il, lsup = (l + 1.0) * il;
  const bool qrow = t >= R_Q0 && t <= R_QLMAX;
  T vq = tconst<T>(1.0), ivq = tconst<T>(1.0);  // v_q and 1/v_q = sqrt(1 + x^2)
  if (qrow) {
    const T aq = a * (cv<T>(cx, F_AMNU) * rcp_nr(cx.q[qi]));
    const T s2 = 1.0 + aq * aq;
    vq = rsqrt_t<T>(s2);
    ivq = s2 * vq;
  }
  switch (t) {
    case R_ETA:
      sink(std::integral_constant<int, C_UETA>{}, 1.0);
      sink(std::integral_constant<int, C_CH>{}, (2.0 * k2) * iaH);
      break;
    case R_DC: {
      const P rc = grhom * Omegac;
      sink(std::integral_constant<int, C_SUP>{}, -1.0);
      sink(std::integral_constant<int, C_UH>{}, -0.5);
      sink(std::integral_constant<int, C_CH>{}, rc * ia * iaH);
      sink(std::integral_constant<int, C_C1>{}, -rc);
      break;
    }
    case R_TC:
      sink(std::integral_constant<int, C_DIAG>{}, -b.aH);
      sink(std::integral_constant<int, C_CETA>{}, grhom * Omegac * ia * hk2);
      break;
    case R_DB: {
      const P rb = grhom * Omegab;
      sink(std::integral_constant<int, C_SUP>{}, -1.0);
      sink(std::integral_constant<int, C_UH>{}, -0.5);
      sink(std::integral_constant<int, C_CH>{}, rb * ia * iaH);
      sink(std::integral_constant<int, C_C1>{}, -rb);
      break;
    }
    case R_TB: {
      const T po = b.pb43 * b.opac;
      sink(std::integral_constant<int, C_SUB>{}, k2 * b.cs2);
      sink(std::integral_constant<int, C_DIAG>{}, -b.aH - po);
      sink(std::integral_constant<int, C_SUP>{}, po);
      sink(std::integral_constant<int, C_CETA>{}, grhom * Omegab * ia * hk2);
      break;
    }
    case R_DG:
      sink(std::integral_constant<int, C_LEAF>{}, -4.0 / 3.0);
      sink(std::integral_constant<int, C_UH>{}, -2.0 / 3.0);
      sink(std::integral_constant<int, C_CH>{}, grhog * ia2 * iaH);
      sink(std::integral_constant<int, C_C1>{}, (-2.0 * grhog) * ia);
      break;
    case R_TG:
      sink(std::integral_constant<int, C_SUB>{}, b.opac);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -0.5 * k2);
      sink(std::integral_constant<int, C_LEAF>{}, 0.25 * k2);
      sink(std::integral_constant<int, C_CETA>{}, (4.0 / 3.0) * grhog * ia2 * hk2);
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
      sink(std::integral_constant<int, C_SUB>{}, 0.4 * k);
      sink(std::integral_constant<int, C_DIAG>{}, -b.opac);
      sink(std::integral_constant<int, C_SUP>{}, -0.6 * k);
      sink(std::integral_constant<int, C_UPI>{}, 0.1 * b.opac);
      sink(std::integral_constant<int, C_CPI>{}, 1.0);
      break;
    case R_GLMAX:
      sink(std::integral_constant<int, C_SUB>{}, k);
      sink(std::integral_constant<int, C_DIAG>{}, -(l + 1.0) * invtau - b.opac);
      break;
    case R_DR: {
      const P rr = grhor * Neff;
      sink(std::integral_constant<int, C_SUP>{}, -4.0 / 3.0);
      sink(std::integral_constant<int, C_UH>{}, -2.0 / 3.0);
      sink(std::integral_constant<int, C_CH>{}, rr * ia2 * iaH);
      sink(std::integral_constant<int, C_C1>{}, (-2.0 * rr) * ia);
      break;
    }
    case R_TR:
      sink(std::integral_constant<int, C_SUB>{}, 0.25 * k2);
      sink(std::integral_constant<int, C_SUP>{}, -0.5 * k2);
      sink(std::integral_constant<int, C_CETA>{}, (4.0 / 3.0) * Neff * grhor * ia2 * hk2);
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
      const P rn = grhor * Nmnu * cx.w[qi];
      sink(std::integral_constant<int, C_SUP>{}, -k * vq);
      sink(std::integral_constant<int, C_UH>{}, cx.dlnf[qi] * (1.0 / 6.0));
      sink(std::integral_constant<int, C_CH>{}, rn * ivq * ia2 * iaH);
      sink(std::integral_constant<int, C_C1>{}, -(rn * (ivq + vq)) * ia);
      break;
    }
    case R_Q1: {
      const T kv = k * vq;
      sink(std::integral_constant<int, C_SUB>{}, kv * (1.0 / 3.0));
      sink(std::integral_constant<int, C_SUP>{}, kv * (-2.0 / 3.0));
      sink(std::integral_constant<int, C_CETA>{}, Nmnu * grhor * k * cx.w[qi] * ia2 * hk2);
      break;
    }
    case R_Q2: {
      const T kv = k * vq;
      sink(std::integral_constant<int, C_SUB>{}, 0.4 * kv);
      sink(std::integral_constant<int, C_SUP>{}, -0.6 * kv);
      sink(std::integral_constant<int, C_UH>{}, cx.dlnf[qi] * (-1.0 / 15.0));
      sink(std::integral_constant<int, C_UETA>{}, -0.4 * cx.dlnf[qi]);
      break;
    }
    case R_QL: {
      const T kv = k * vq;
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
        sink(std::integral_constant<int, C_SUP>{}, -opw - 9.0 * opw * (cs2Q - b.ca2) * b.aH * b.aH * ik2);
        sink(std::integral_constant<int, C_UH>{}, -0.5 * opw);
        T rq = grhom * OmegaDE * b.rhoQ * (a * a);
        sink(std::integral_constant<int, C_CH>{}, rq * iaH);
        sink(std::integral_constant<int, C_C1>{}, -a * (rq + 3.0 * cs2Q * rq));
      }
      break;
    case R_TQ:
      if (!cx.lambda) {
        const P cs2Q = cs2DE;
        T opw = 1.0 + b.w;
        sink(std::integral_constant<int, C_SUB>{}, cs2Q / opw * k2);
        sink(std::integral_constant<int, C_DIAG>{}, -(1.0 - 3.0 * cs2Q) * b.aH);
        T rq = grhom * OmegaDE * b.rhoQ * (a * a);
        sink(std::integral_constant<int, C_CETA>{}, opw * rq * hk2);
        sink(std::integral_constant<int, C_C1>{}, -a * (3.0 * (cs2Q - b.ca2) * (3.0 * b.aH * ik2) * opw * rq));
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
  // (accumulating rather than assigning keeps the compiler from merging the
  // cases' stores into one store at a variable index, which would move v to
  // the stack)
  row_visit<T>(d, b, cx, invtau, [&](auto f, const auto &x) __attribute__((always_inline)) {
    v[decltype(f)::value] = v[decltype(f)::value] + x;
  });
  Coef<T> r;
  r.sub = v[C_SUB]; r.diag = v[C_DIAG]; r.sup = v[C_SUP];
  r.uh = v[C_UH]; r.ueta = v[C_UETA]; r.upi = v[C_UPI];
  r.ch = v[C_CH]; r.ceta = v[C_CETA]; r.cpi = v[C_CPI]; r.c1 = v[C_C1]; r.leaf = v[C_LEAF];
  return r;
}

// row_coef organised for the wave instead of as a divergent switch, with the
// same formulas in the same operation order (the host emulation compares the
// two): the chain rows (hierarchies, any slot) by per-lane selects among their
// formulas; with slot0, the fixed-lane rows of slot 0 as wave-uniform values
// selected by lane.
template <typename T>
__device__ __forceinline__ Coef<T> row_coef_wave(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau,
                                                 bool slot0) {
  using P = ParT<T>;
  const T zero = tconst<T>(0.0);
  auto c = [](double v) { return tconst<T>(v); };
  const P grhom = cv<T>(cx, F_GRHOM), grhog = cv<T>(cx, F_GRHOG), grhor = cv<T>(cx, F_GRHOR);
  const P Neff = cv<T>(cx, F_NEFF), Nmnu = cv<T>(cx, F_NMNU);
  const double k = cx.k, k2 = k * k;
  const double ik2 = 1.0 / k2, hk2 = 0.5 * ik2;
  const uint32_t t = rd_type(d);
  const double l = (double)rd_l(d);
  const bool F2 = t == R_F2, FL = t == R_FL, FLM = t == R_FLMAX, G0 = t == R_G0, GL = t == R_GL, G2 = t == R_G2,
             GLM = t == R_GLMAX, DR = t == R_DR, TR = t == R_TR, FR2 = t == R_FR2, FRL = t == R_FRL,
             FRLM = t == R_FRLMAX, Q0 = t == R_Q0, Q1 = t == R_Q1, Q2 = t == R_Q2, QL = t == R_QL,
             QLM = t == R_QLMAX;
  const bool qrow = t >= R_Q0 && t <= R_QLMAX;
  const int qi = qrow ? (int)rd_q(d) : 0;
  const T ia = b.ia, ia2 = b.ia2, iaH = b.iaH;
  const double il = rcp_nr(2.0 * l + 1.0);
  const double lsub = l * il, lsup = (l + 1.0) * il;
  // v_q and 1/v_q of the row's bin (1 on other rows); k v_q multiplies the
  // couplings of every chain (k v_q = k exactly where v_q = 1)
  T vq = c(1.0), ivq = c(1.0);
  {
    const T aq = b.a * (cv<T>(cx, F_AMNU) * rcp_nr(cx.q[qi]));
    const T s2 = 1.0 + aq * aq;
    const T v = rsqrt_t<T>(s2);
    vq = qrow ? v : c(1.0);
    ivq = qrow ? s2 * v : c(1.0);
  }
  const T kv = k * vq;
  const double wq = cx.w[qi], dl = cx.dlnf[qi];
  Coef<T> r;
  // sub-diagonal
  double cs = 0.0;
  cs = (FL || GL || FRL || QL) ? lsub : cs;
  cs = (FLM || GLM || FRLM || QLM) ? 1.0 : cs;
  cs = (G2 || Q2) ? 0.4 : cs;
  cs = Q1 ? 1.0 / 3.0 : cs;
  r.sub = kv * cs;
  r.sub = (F2 || FR2) ? c(8.0 / 15.0) : r.sub;
  r.sub = TR ? c(0.25 * k2) : r.sub;
  // diagonal: opacity of the photon rows, free streaming out of the last multipole
  const bool photon = F2 || FL || FLM || G0 || GL || G2 || GLM;
  r.diag = photon ? -b.opac : zero;
  r.diag = (FLM || GLM) ? -(l + 1.0) * invtau - b.opac : r.diag;
  r.diag = (FRLM || QLM) ? c(-(l + 1.0) * invtau) : r.diag;
  // super-diagonal
  double cp = 0.0;
  cp = (F2 || G2 || FR2 || Q2) ? 0.6 : cp;
  cp = (FL || GL || FRL || QL) ? lsup : cp;
  cp = (G0 || Q0) ? 1.0 : cp;
  cp = Q1 ? 2.0 / 3.0 : cp;
  r.sup = cp != 0.0 ? -(kv * cp) : zero;
  r.sup = DR ? c(-4.0 / 3.0) : r.sup;
  r.sup = TR ? c(-0.5 * k2) : r.sup;
  // where h', eta', Pi enter
  r.uh = (F2 || FR2) ? c(4.0 / 15.0) : zero;
  r.uh = DR ? c(-2.0 / 3.0) : r.uh;
  r.uh = Q0 ? c(dl * (1.0 / 6.0)) : r.uh;
  r.uh = Q2 ? c(dl * (-1.0 / 15.0)) : r.uh;
  r.ueta = (F2 || FR2) ? c(8.0 / 5.0) : (Q2 ? c(-0.4 * dl) : zero);
  r.upi = (F2 || G2) ? 0.1 * b.opac : (G0 ? 0.5 * b.opac : zero);
  // contributions to h', eta', Pi and f_1
  const P rr = grhor * Neff;
  const P rn = grhor * Nmnu * wq;
  r.ch = DR ? rr * ia2 * iaH : (Q0 ? rn * ivq * ia2 * iaH : zero);
  r.ceta = TR ? (4.0 / 3.0) * Neff * grhor * ia2 * hk2 : (Q1 ? Nmnu * grhor * k * wq * ia2 * hk2 : zero);
  r.cpi = (F2 || G0 || G2) ? c(1.0) : zero;
  r.c1 = DR ? (-2.0 * rr) * ia : (Q0 ? -(rn * (ivq + vq)) * ia : zero);
  r.leaf = zero;
  if (slot0) {
    const int lane = lane_id();
    const P Omegam = cv<T>(cx, F_OMEGAM), Omegab = cv<T>(cx, F_OMEGAB);
    const P Omegac = Omegam - Omegab;
    const T a = b.a;
    // eta
    r.ueta = lane == L_ETA ? c(1.0) : r.ueta;
    r.ch = lane == L_ETA ? (2.0 * k2) * iaH : r.ch;
    // delta_c, delta_b
    const P rc = grhom * Omegac, rb = grhom * Omegab;
    const bool dcb = lane == L_DC || lane == L_DB;
    const P rcb = lane == L_DC ? rc : rb;
    r.sup = dcb ? c(-1.0) : r.sup;
    r.uh = dcb ? c(-0.5) : r.uh;
    r.ch = dcb ? rcb * ia * iaH : r.ch;
    r.c1 = dcb ? zero - rcb : r.c1;
    // theta_c
    r.diag = lane == L_TC ? -b.aH : r.diag;
    r.ceta = lane == L_TC ? grhom * Omegac * ia * hk2 : r.ceta;
    // theta_b
    {
      const T po = b.pb43 * b.opac;
      const bool on = lane == L_TB;
      r.sub = on ? k2 * b.cs2 : r.sub;
      r.diag = on ? -b.aH - po : r.diag;
      r.sup = on ? po : r.sup;
      r.ceta = on ? grhom * Omegab * ia * hk2 : r.ceta;
    }
    // delta_gamma
    {
      const bool on = lane == L_DG;
      r.leaf = on ? c(-4.0 / 3.0) : r.leaf;
      r.uh = on ? c(-2.0 / 3.0) : r.uh;
      r.ch = on ? grhog * ia2 * iaH : r.ch;
      r.c1 = on ? (-2.0 * grhog) * ia : r.c1;
    }
    // theta_gamma
    {
      const bool on = lane == L_TG;
      r.sub = on ? b.opac : r.sub;
      r.diag = on ? -b.opac : r.diag;
      r.sup = on ? c(-0.5 * k2) : r.sup;
      r.leaf = on ? c(0.25 * k2) : r.leaf;
      r.ceta = on ? (4.0 / 3.0) * grhog * ia2 * hk2 : r.ceta;
    }
    // dark-energy fluid (frozen for a cosmological constant)
    if (!cx.lambda) {
      const P OmegaDE = cv<T>(cx, F_OMEGADE), cs2Q = cv<T>(cx, F_CS2DE);
      const T opw = 1.0 + b.w;
      const T rq = grhom * OmegaDE * b.rhoQ * (a * a);
      const bool dq = lane == L_DQ, tq = lane == L_TQ;
      r.diag = dq ? -3.0 * (cs2Q - b.w) * b.aH : r.diag;
      r.sup = dq ? -opw - 9.0 * opw * (cs2Q - b.ca2) * b.aH * b.aH * ik2 : r.sup;
      r.uh = dq ? -0.5 * opw : r.uh;
      r.ch = dq ? rq * iaH : r.ch;
      r.c1 = dq ? -a * (rq + 3.0 * cs2Q * rq) : r.c1;
      r.sub = tq ? cs2Q / opw * k2 : r.sub;
      r.diag = tq ? -(1.0 - 3.0 * cs2Q) * b.aH : r.diag;
      r.ceta = tq ? opw * rq * hk2 : r.ceta;
      r.c1 = tq ? -a * (3.0 * (cs2Q - b.ca2) * (3.0 * b.aH * ik2) * opw * rq) : r.c1;
    }
  }
  return r;
}

// row_coef of slot s: the wave form in plain doubles; dual arithmetic (the
// factorisation's a-column and the reverse step's a-column) keeps the switch.
// With the wave form also in the factorisation's dual right-hand side, the
// fully inlined forward kernel ended 41 % of the headline modes at max_steps on
// the GPU, although the host emulation finds the dual wave form equal to the
// switch in every field; the cause was not found.
template <typename T>
__device__ __forceinline__ Coef<T> row_coef_slot(uint32_t d, const Bg<T> &b, const Ctx &cx, double invtau,
                                                 bool slot0) {
  if constexpr (std::is_same<T, double>::value) return row_coef_wave<T>(d, b, cx, invtau, slot0);
  else return row_coef<T>(d, b, cx, invtau);
}

}  // namespace deb
