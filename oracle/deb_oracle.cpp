// DISCO-EB perturbation hot path: CPU restatement used as the ORACLE.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library; the product path never
// does.  It restates, in plain FP64 C++, what the reference computes:
//
//   model_synchronous          src/discoeb/perturbations.py:84-371
//   nu_perturb                 src/discoeb/perturbations.py:22-50
//   get_aprimeoa               src/discoeb/background.py:100-122
//   spline evaluate            src/discoeb/spline_interpolation.py:130-153
//   convert_to_output_vars     src/discoeb/perturbations.py:374-523
//   adiabatic_ics_one_mode     src/discoeb/perturbations.py:526-627
//   determine_starting_time    src/discoeb/perturbations.py:630-681 (+ util.py:365-396)
//   evolve_one_mode            src/discoeb/perturbations.py:728-781
//   evolve_modes_batched       src/discoeb/perturbations.py:786-922 (one step
//                              sequence per group: start time = group minimum,
//                              norm = rms_norm_filtered_batched :690-694)
//   rms_norm_filtered          src/discoeb/perturbations.py:701-705
//   Rodas5Transformed.step     src/discoeb/ode_integrators_stiff.py:622-834
//   get_power                  src/discoeb/perturbations.py:1101-1123
//   diffrax loop semantics     PIDController (p, i, d, factormin only on
//                              rejection, safety 0.9, error order 5), dt0,
//                              _clip_to_end (1e-10), SaveAt linear
//                              interpolation, max_steps -> status.
//
// The Jacobian is exact (forward-mode duals, like jax.jacfwd) and the linear
// algebra is a dense partial-pivot LU (like jax.scipy.linalg.lu_factor).
//
// Gradients (the reference differentiates the whole pipeline with jax.jacfwd,
// nb_minimal_example.ipynb:297): every function is templated on the scalar
// type, so the same code runs on nested dual numbers.  The tangent solve
// (oracle_evolve_tangent) replays the accepted steps of the plain solve with
// the step-size controller's decisions held fixed: step n starts at
// t_n = t0 + s_n (t1 - t0) with the fractions s_n of the plain solve, while
// t0 (start time), t1 (last output time), the initial state, every cosmology
// scalar and every spline knot carry the tangent.  The result is the exact
// directional derivative of the discrete solution.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "../include/discoeb_hip.h"

namespace {

// ---------------------------------------------------------------------------
// nestable forward-mode dual numbers: D<double> carries one tangent,
// D<D<double>> a derivative whose value and slope both carry a tangent
template <class B>
struct D {
  B v, d;
  D() : v(0.0), d(0.0) {}
  D(const B &v_) : v(v_), d(0.0) {}
  D(const B &v_, const B &d_) : v(v_), d(d_) {}
  template <class X, class = std::enable_if_t<std::is_arithmetic_v<X> && !std::is_same_v<X, B>>>
  D(X x) : v(B(double(x))), d(0.0) {}
};

template <class S, class B>
constexpr bool kScalarOf = std::is_same_v<S, double> || std::is_same_v<S, int> || std::is_same_v<S, B>;

template <class B>
D<B> operator+(const D<B> &a, const D<B> &b) { return {a.v + b.v, a.d + b.d}; }
template <class B>
D<B> operator-(const D<B> &a, const D<B> &b) { return {a.v - b.v, a.d - b.d}; }
template <class B>
D<B> operator-(const D<B> &a) { return {-a.v, -a.d}; }
template <class B>
D<B> operator*(const D<B> &a, const D<B> &b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class B>
D<B> operator/(const D<B> &a, const D<B> &b) {
  B q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator+(const D<B> &a, const S &b) { return {a.v + b, a.d}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator+(const S &b, const D<B> &a) { return {b + a.v, a.d}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator-(const D<B> &a, const S &b) { return {a.v - b, a.d}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator-(const S &b, const D<B> &a) { return {b - a.v, -a.d}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator*(const D<B> &a, const S &b) { return {a.v * b, a.d * b}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator*(const S &b, const D<B> &a) { return {b * a.v, b * a.d}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator/(const D<B> &a, const S &b) { return {a.v / b, a.d / b}; }
template <class B, class S, class = std::enable_if_t<kScalarOf<S, B>>>
D<B> operator/(const S &b, const D<B> &a) {
  B q = b / a.v;
  return {q, -q * a.d / a.v};
}
using std::exp;
using std::log;
using std::pow;
using std::sqrt;
template <class B>
D<B> exp(const D<B> &a) {
  B e = exp(a.v);
  return {e, e * a.d};
}
template <class B>
D<B> log(const D<B> &a) { return {log(a.v), a.d / a.v}; }
template <class B>
D<B> sqrt(const D<B> &a) {
  B s = sqrt(a.v);
  return {s, 0.5 * a.d / s};
}
template <class B>
D<B> pow(const D<B> &a, double p) { return {pow(a.v, p), p * pow(a.v, p - 1.0) * a.d}; }
inline double val(double x) { return x; }
template <class B>
double val(const D<B> &x) { return val(x.v); }

using Dd = D<double>;

// ---------------------------------------------------------------------------
// natural cubic spline evaluation: interval = clip(searchsorted_left(x, v)-1, 0, n-2)
template <class P>
struct SplineT {
  const P *x, *y, *S;
  int n;
};

template <class P>
int interval(const SplineT<P> &s, double v) {
  // number of knots strictly below v (left-sided searchsorted)
  int lo = 0, hi = s.n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (val(s.x[mid]) < v) lo = mid + 1; else hi = mid;
  }
  int i = lo - 1;
  if (i < 0) i = 0;
  if (i > s.n - 2) i = s.n - 2;
  return i;
}

template <class T, class P>
T spline_eval(const SplineT<P> &s, const T &xv) {
  int i = interval(s, val(xv));
  P h = s.x[i + 1] - s.x[i];
  T t = (xv - s.x[i]) / h;
  T A = 1.0 - t;
  T B = t;
  return A * s.y[i] + B * s.y[i + 1] + ((A * A * A - A) * s.S[i] + (B * B * B - B) * s.S[i + 1]) * ((h * h) / 6.0);
}

template <class P>
struct CosmoT {
  P grhom, grhog, grhor, Omegam, Omegab, OmegaDE, Omegak, Neff, Nmnu, amnu, w_DE_0, w_DE_a, cs2_DE, YHe, H0,
      taumin;
};

CosmoT<double> cosmo_of(const deb_cosmo_t &c) {
  return {c.grhom, c.grhog, c.grhor, c.Omegam, c.Omegab, c.OmegaDE, c.Omegak, c.Neff, c.Nmnu, c.amnu,
          c.w_DE_0, c.w_DE_a, c.cs2_DE, c.YHe, c.H0, c.taumin};
}

CosmoT<Dd> cosmo_of(const deb_cosmo_t &c, const deb_cosmo_t &t) {
  return {{c.grhom, t.grhom}, {c.grhog, t.grhog}, {c.grhor, t.grhor}, {c.Omegam, t.Omegam},
          {c.Omegab, t.Omegab}, {c.OmegaDE, t.OmegaDE}, {c.Omegak, t.Omegak}, {c.Neff, t.Neff},
          {c.Nmnu, t.Nmnu}, {c.amnu, t.amnu}, {c.w_DE_0, t.w_DE_0}, {c.w_DE_a, t.w_DE_a},
          {c.cs2_DE, t.cs2_DE}, {c.YHe, t.YHe}, {c.H0, t.H0}, {c.taumin, t.taumin}};
}

template <class P>
struct ModelT {
  CosmoT<P> c;
  SplineT<P> sp[DEB_NSPLINE];
  double k;
  int lmaxg, lmaxgp, lmaxr, lmaxnu, nq;
  double q[DEB_MAXNQ], w[DEB_MAXNQ];
  bool lambda;  // w0 == -1 and wa == 0
  int nvar() const { return 12 + lmaxg + lmaxgp + lmaxr + nq * (lmaxnu + 1); }
};
using Model = ModelT<double>;

// rho_Q(a) = a^{-3(1+w0+wa)} exp(3 (a-1) wa), with the exponent's tangent
template <class T, class P>
T rho_de(const ModelT<P> &m, const T &a) {
  const CosmoT<P> &c = m.c;
  const double e0 = -3.0 * val(1.0 + c.w_DE_0 + c.w_DE_a);
  T r = pow(a, e0) * exp(3.0 * (a - 1.0) * c.w_DE_a);
  if constexpr (!std::is_same_v<P, double>) r = r * exp((-3.0 * (1.0 + c.w_DE_0 + c.w_DE_a) - e0) * log(a));
  return r;
}

template <class T, class P>
T aprimeoa(const ModelT<P> &m, const T &a) {
  const CosmoT<P> &c = m.c;
  T rhonu = exp(spline_eval(m.sp[DEB_SPL_LOGRHONU_LOGA], log(a)));
  T rhoQ = rho_de(m, a);
  T grho = c.grhom * c.Omegam / a + (c.grhog + c.grhor * (c.Neff + c.Nmnu * rhonu)) / (a * a) +
           c.grhom * c.OmegaDE * rhoQ * a * a + c.grhom * c.Omegak;
  return sqrt(grho / 3.0);
}

// model_synchronous: f = dy/dtau
template <class T, class P>
void rhs(const ModelT<P> &m, const T &tau, const T *y, T *f) {
  const CosmoT<P> &c = m.c;
  const int lg = m.lmaxg, lgp = m.lmaxgp, lr = m.lmaxr, lnu = m.lmaxnu, nq = m.nq;
  const int n = m.nvar();
  const double k = m.k, k2 = k * k;
  const P Omegac = c.Omegam - c.Omegab;
  const int iq0 = 10 + lg + lgp + lr;
  const int iq1 = iq0 + nq, iq2 = iq1 + nq, iq3 = iq2 + nq;
  T a = y[0];
  T loga = log(a);
  T eta = y[2], deltac = y[3], thetac = y[4], deltab = y[5], thetab = y[6];
  T deltag = y[7], thetag = y[8];
  T deltar = y[9 + lg + lgp], thetar = y[10 + lg + lgp];
  T deltaq = y[n - 2], thetaq = y[n - 1];
  T cs2 = spline_eval(m.sp[DEB_SPL_CS2A_LOGA], loga) / a;
  T xe = spline_eval(m.sp[DEB_SPL_XE_LOGA], loga);
  T photbar = c.grhog / (c.grhom * c.Omegab * a);
  T pb43 = 4.0 / 3.0 * photbar;
  const P cs2_Q = c.cs2_DE;
  T w_Q = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  T rho_Q = rho_de(m, a);
  T rppQ = (1.0 + w_Q) * rho_Q * c.grhom * c.OmegaDE * thetaq * a * a;
  T aH = aprimeoa(m, a);
  T w_Q_prime = -c.w_DE_a * aH * a;
  T ca2_Q = w_Q - w_Q_prime / 3.0 / ((1.0 + w_Q) + 1e-6) / aH;
  const P akthom = 2.3038921003709498e-9 * (1.0 - c.YHe) * c.Omegab * c.H0 * c.H0;
  T opac = xe * akthom / (a * a);
  f[0] = aH * a;
  // nu_perturb
  T drhonu = 0.0, dpnu = 0.0, fnu = 0.0, shearnu = 0.0;
  for (int i = 0; i < nq; ++i) {
    T aq = a * c.amnu / m.q[i];
    T v = 1.0 / sqrt(1.0 + aq * aq);
    drhonu = drhonu + m.w[i] * y[iq0 + i] / v;
    dpnu = dpnu + m.w[i] * y[iq0 + i] * v;
    fnu = fnu + m.w[i] * y[iq1 + i];
    shearnu = shearnu + m.w[i] * y[iq2 + i] * v;
  }
  dpnu = dpnu / 3.0;
  shearnu = shearnu * 2.0 / 3.0;
  T a2 = a * a;
  T dgrho = c.grhom * (Omegac * deltac + c.Omegab * deltab) / a +
            (c.grhog * deltag + c.grhor * (c.Neff * deltar + c.Nmnu * drhonu)) / a2 +
            c.grhom * c.OmegaDE * deltaq * rho_Q * a2;
  T dgpres = (c.grhog * deltag + c.grhor * c.Neff * deltar) / a2 / 3.0 + c.grhor * c.Nmnu * dpnu / a2 +
             (cs2_Q * c.grhom * c.OmegaDE * deltaq * rho_Q * a2 + (cs2_Q - ca2_Q) * (3.0 * aH * rppQ / k2));
  T dgtheta = c.grhom * (Omegac * thetac + c.Omegab * thetab) / a +
              4.0 / 3.0 * (c.grhog * thetag + c.Neff * c.grhor * thetar) / a2 + c.Nmnu * c.grhor * k * fnu / a2 + rppQ;
  f[1] = -(dgrho + 3.0 * dgpres) * a;
  T hprime = (2.0 * k2 * eta + dgrho) / aH;
  T etaprime = 0.5 * dgtheta / k2;
  T alpha = (hprime + 6.0 * etaprime) / 2.0 / k2;
  f[2] = etaprime;
  f[3] = -thetac - 0.5 * hprime;
  f[4] = -aH * thetac;
  f[5] = -thetab - 0.5 * hprime;
  f[6] = -aH * thetab + k2 * cs2 * deltab + pb43 * opac * (thetag - thetab);
  const int ig = 7, igp = 8 + lg;
  T polter = y[ig + 2] + y[igp + 0] + y[igp + 2];
  f[ig + 0] = 4.0 / 3.0 * (-thetag - 0.5 * hprime);
  f[ig + 1] = k2 * (0.25 * deltag - y[ig + 2] / 2.0) - opac * (thetag - thetab);
  f[ig + 2] = 8.0 / 15.0 * (thetag + k2 * alpha) - 3.0 / 5.0 * k * y[ig + 3] - opac * (y[ig + 2] - 0.1 * polter);
  for (int l = 3; l < lg; ++l)
    f[ig + l] = k / (2.0 * l + 1.0) * (l * y[ig + l - 1] - (l + 1.0) * y[ig + l + 1]) - opac * y[ig + l];
  f[ig + lg] = k * y[ig + lg - 1] - (lg + 1.0) / tau * y[ig + lg] - opac * y[ig + lg];
  for (int l = 0; l < lgp; ++l) {
    T ym = (l > 0) ? y[igp + l - 1] : T(0.0);
    f[igp + l] = k / (2.0 * l + 1.0) * (l * ym - (l + 1.0) * y[igp + l + 1]) - opac * y[igp + l];
  }
  f[igp + 0] = f[igp + 0] + opac * polter / 2.0;
  f[igp + 2] = f[igp + 2] + opac * polter / 10.0;
  f[igp + lgp] = k * y[igp + lgp - 1] - (lgp + 1.0) / tau * y[igp + lgp] - opac * y[igp + lgp];
  const int ir = 9 + lg + lgp;
  f[ir + 0] = 4.0 / 3.0 * (-thetar - 0.5 * hprime);
  f[ir + 1] = k2 * (0.25 * deltar - y[ir + 2] / 2.0);
  f[ir + 2] = 8.0 / 15.0 * (thetar + k2 * alpha) - 0.6 * k * y[ir + 3];
  for (int l = 3; l < lr; ++l) f[ir + l] = k / (2.0 * l + 1.0) * (l * y[ir + l - 1] - (l + 1.0) * y[ir + l + 1]);
  f[ir + lr] = k * y[ir + lr - 1] - (lr + 1.0) / tau * y[ir + lr];
  for (int i = 0; i < nq; ++i) {
    double qi = m.q[i];
    T aq = a * c.amnu / qi;
    T v = 1.0 / sqrt(1.0 + aq * aq);
    double dlfdlq = -qi / (1.0 + std::exp(-qi));
    f[iq0 + i] = -k * v * y[iq1 + i] + hprime * dlfdlq / 6.0;
    f[iq1 + i] = k * v * (y[iq0 + i] - 2.0 * y[iq2 + i]) / 3.0;
    f[iq2 + i] = k * v * (2.0 * y[iq1 + i] - 3.0 * y[iq3 + i]) / 5.0 - (hprime / 15.0 + 2.0 / 5.0 * etaprime) * dlfdlq;
    for (int l = 3; l < lnu; ++l)
      f[iq0 + l * nq + i] = k * v / (2.0 * l + 1.0) *
                            (l * y[iq0 + (l - 1) * nq + i] - (l + 1.0) * y[iq0 + (l + 1) * nq + i]);
    f[iq0 + lnu * nq + i] = k * v * y[iq0 + (lnu - 1) * nq + i] - (lnu + 1.0) / tau * y[iq0 + lnu * nq + i];
  }
  f[n - 2] = -(1.0 + w_Q) * (thetaq + 0.5 * hprime) - 3.0 * (cs2_Q - w_Q) * aH * deltaq -
             9.0 * (1.0 + w_Q) * (cs2_Q - ca2_Q) * aH * aH / k2 * thetaq;
  f[n - 1] = -(1.0 - 3.0 * cs2_Q) * aH * thetaq + cs2_Q / (1.0 + w_Q) * k2 * deltaq;
  if (m.lambda) {
    // cosmological constant: the fluid equations are singular at w = -1
    // (inf * 0 in theta_Q'); their w -> -1 limit keeps delta_Q = theta_Q = 0
    f[n - 2] = T(0.0);
    f[n - 1] = T(0.0);
  }
}

template <class T, class P>
void outputs(const ModelT<P> &m, const T *y, T *out) {
  const CosmoT<P> &c = m.c;
  const int lg = m.lmaxg, lgp = m.lmaxgp, lr = m.lmaxr, nq = m.nq;
  const int n = m.nvar();
  const double k = m.k, k2 = k * k;
  const P Omegac = c.Omegam - c.Omegab;
  const int iq0 = 10 + lg + lgp + lr, iq1 = iq0 + nq;
  T a = y[0], eta = y[2], deltac = y[3], thetac = y[4], deltab = y[5], thetab = y[6];
  T deltag = y[7], thetag = y[8];
  T deltar = y[9 + lg + lgp], thetar = y[10 + lg + lgp];
  T rhonu = exp(spline_eval(m.sp[DEB_SPL_LOGRHONU_LOGA], log(a)));
  T pnu = exp(spline_eval(m.sp[DEB_SPL_LOGPNU_LOGA], log(a)));
  T drhonu = 0.0, fnu = 0.0;
  for (int i = 0; i < nq; ++i) {
    T aq = a * c.amnu / m.q[i];
    T v = 1.0 / sqrt(1.0 + aq * aq);
    drhonu = drhonu + m.w[i] * y[iq0 + i] / v;
    fnu = fnu + m.w[i] * y[iq1 + i];
  }
  T deltanu = drhonu / rhonu;
  T thetanu = k * fnu / (rhonu + pnu);
  T deltaq = y[n - 2], thetaq = y[n - 1];
  T w_Q = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  T rho_Q = rho_de(m, a);
  T rppQ = (1.0 + w_Q) * rho_Q * c.grhom * c.OmegaDE * thetaq * a * a;
  T a2 = a * a;
  T grho = c.grhom * c.Omegam / a + (c.grhog + c.grhor * (c.Neff + c.Nmnu * rhonu)) / a2 +
           c.grhom * c.OmegaDE * rho_Q * a2 + c.grhom * c.Omegak;
  T aH = sqrt(grho / 3.0);
  T dgrho = c.grhom * (Omegac * deltac + c.Omegab * deltab) / a +
            (c.grhog * deltag + c.grhor * (c.Neff * deltar + c.Nmnu * drhonu)) / a2 +
            c.grhom * c.OmegaDE * deltaq * rho_Q * a2;
  T dgtheta = c.grhom * (Omegac * thetac + c.Omegab * thetab) / a +
              4.0 / 3.0 * (c.grhog * thetag + c.Neff * c.grhor * thetar) / a2 + c.Nmnu * c.grhor * k * fnu / a2 + rppQ;
  T hprime = (2.0 * k2 * eta + dgrho) / aH;
  T etaprime = 0.5 * dgtheta / k2;
  T alpha = (hprime + 6.0 * etaprime) / 2.0 / k2;
  T mden = c.grhom * c.Omegam / a + c.grhor * c.Nmnu * rhonu / a2;
  T deltam = (c.grhom * (Omegac * deltac + c.Omegab * deltab) / a + c.grhor * c.Nmnu * drhonu / a2) / mden;
  T thetam = (c.grhom * (Omegac * thetac + c.Omegab * thetab) / a + c.Nmnu * c.grhor * k * fnu / a2) / (3.0 * mden);
  T deltabc = (c.grhom * (Omegac * deltac + c.Omegab * deltab) / a) / (c.grhom * c.Omegam / a);
  T thetabc = (c.grhom * (Omegac * thetac + c.Omegab * thetab) / a) / (3.0 * (c.grhom * c.Omegam / a) / a2);
  thetam = thetam + alpha * k2;
  thetabc = thetabc + alpha * k2;
  const T o[20] = {eta, etaprime, hprime, alpha, deltam, thetam / aH, deltabc, thetabc / aH,
                   deltac, thetac / aH, deltab, thetab / aH, deltag, thetag / aH, deltar, thetar / aH,
                   deltanu, thetanu / aH, deltaq, thetaq / aH};
  for (int i = 0; i < 20; ++i) out[i] = o[i];
}

template <class P>
void ics(const ModelT<P> &m, const P &tau, P *y) {
  const CosmoT<P> &c = m.c;
  const int lg = m.lmaxg, lgp = m.lmaxgp, lr = m.lmaxr, nq = m.nq;
  const int n = m.nvar();
  const int iq0 = 10 + lg + lgp + lr, iq1 = iq0 + nq, iq2 = iq1 + nq;
  const double k = m.k;
  for (int i = 0; i < n; ++i) y[i] = P(0.0);
  P a = spline_eval(m.sp[DEB_SPL_A_TAU], tau);
  P rnu = exp(spline_eval(m.sp[DEB_SPL_LOGRHONU_LOGA], log(a)));
  P rhom = c.grhom * c.Omegam / (a * a * a);
  P rhor = (c.grhog + c.grhor * (c.Neff + c.Nmnu * rnu)) / pow(a, 4.0);
  P rhonu = c.grhor * (c.Neff + c.Nmnu * rnu) / pow(a, 4.0);
  P fracb = c.Omegab / c.Omegam;
  P fracnu = rhonu / rhor;
  P om = a * rhom / sqrt(rhor);
  const double C = -1.0, s2 = 1.0;
  P kt = k * tau;
  P deltag = -(kt * kt) / 3.0 * (1.0 - om * tau / 5.0) * C * s2;
  P thetag = -(kt * kt * kt) / tau / 36.0 * (1.0 - 3.0 * (1.0 + 5.0 * fracb - fracnu) / 20.0 / (1.0 - fracnu) * om * tau) * C * s2;
  P deltab = 0.75 * deltag, thetab = thetag, deltac = 0.75 * deltag, thetac = 0.0;
  P deltar = deltag;
  P thetar = -(kt * kt * kt * kt) / tau / 36.0 / (4.0 * fracnu + 15.0) *
             (4.0 * fracnu + 11.0 + 12.0 - 3.0 * (8.0 * fracnu * fracnu + 50.0 * fracnu + 275.0) / 20.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  P shearr = (kt * kt) / (45.0 + 12.0 * fracnu) * (3.0 * s2 - 1.0) * (1.0 + (4.0 * fracnu - 5.0) / 4.0 / (2.0 * fracnu + 15.0) * tau * om) * C;
  const P cs2_Q = c.cs2_DE;
  P w_Q = c.w_DE_0 + c.w_DE_a * (1.0 - a);
  P deltaq = (kt * kt) / 4.0 * (1.0 + w_Q) * (4.0 - 3.0 * cs2_Q) / (4.0 - 6.0 * w_Q + 3.0 * cs2_Q) * C * s2;
  P thetaq = (kt * kt * kt * kt) / tau / 4.0 * cs2_Q / (4.0 - 6.0 * w_Q + 3.0 * cs2_Q) * C * s2;
  P eta = C * (1.0 - (kt * kt) / 12.0 / (15.0 + 4.0 * fracnu) *
                         (5.0 + 4.0 * s2 * fracnu - (16.0 * fracnu * fracnu + 280.0 * fracnu + 325.0) / 10.0 / (2.0 * fracnu + 15.0) * tau * om));
  y[0] = a;
  y[1] = 0.0;
  y[2] = eta;
  y[3] = deltac;
  y[4] = thetac;
  y[5] = deltab;
  y[6] = thetab;
  y[7] = deltag;
  y[8] = thetag;
  y[9 + lg + lgp] = deltar;
  y[10 + lg + lgp] = thetar;
  y[11 + lg + lgp] = shearr * 2.0;
  for (int i = 0; i < nq; ++i) {
    double qi = m.q[i];
    P aq = a * c.amnu / qi;
    P v = 1.0 / sqrt(1.0 + aq * aq);
    double dlfdlq = -qi / (1.0 + std::exp(-qi));
    y[iq0 + i] = -0.25 * dlfdlq * deltar;
    y[iq1 + i] = -dlfdlq * thetar / v / k / 3.0;
    y[iq2 + i] = -0.5 * dlfdlq * shearr;
  }
  y[n - 2] = deltaq;
  y[n - 1] = thetaq;
  if (m.lambda) {
    y[n - 2] = 0.0;
    y[n - 1] = 0.0;
  }
}

template <class P>
P start_time(const ModelT<P> &m, const P &tau_out_min) {
  const CosmoT<P> &c = m.c;
  P tau0 = c.taumin;
  P tau1 = spline_eval(m.sp[DEB_SPL_TAU_A], P(0.1));
  double tau_k = 1.0 / m.k;
  P akthom = 2.3048e-9 * (1.0 - c.YHe) * c.Omegab * c.H0 * c.H0;
  // the bisections below only compare signs, so the tangent of the root
  // flows through the bracket end points
  auto small = [&](const P &lt) {
    P tau = exp(lt);
    P xe = spline_eval(m.sp[DEB_SPL_XE_TAU], tau);
    P a = spline_eval(m.sp[DEB_SPL_A_TAU], tau);
    P opac = xe * akthom / (a * a);
    P aH = aprimeoa(m, a);
    return val((1.0 / opac) / (1.0 / aH) / 0.0004 - 1.0);
  };
  auto large = [&](const P &lt) {
    P tau = exp(lt);
    P a = spline_eval(m.sp[DEB_SPL_A_TAU], tau);
    P aH = aprimeoa(m, a);
    return val((1.0 / aH) / tau_k / 0.07 - 1.0);
  };
  auto bisect = [&](auto fn) {
    P xl = log(tau0), xr = log(tau1);
    for (int i = 0; i < 7; ++i) {
      P xm = 0.5 * (xl + xr);
      if (fn(xm) * fn(xl) > 0) xl = xm; else xr = xm;
    }
    return P(0.5 * (xl + xr));
  };
  P lt_large = bisect(large);
  P lt_small = bisect(small);
  P ts = exp(val(lt_small) < val(lt_large) ? lt_small : lt_large);
  P tmin = val(tau_out_min) < val(ts) ? tau_out_min : ts;
  return 0.99 * tmin;
}

// ---------------------------------------------------------------------------
// Rodas5 (W-form) constants, ode_integrators_stiff.py:622-687
namespace r5 {
const double gamma = 0.19;
const double a21 = 2.0, a31 = 3.040894194418781, a32 = 1.041747909077569, a41 = 2.576417536461461,
             a42 = 1.622083060776640, a43 = -0.9089668560264532, a51 = 2.760842080225597, a52 = 1.446624659844071,
             a53 = -0.3036980084553738, a54 = 0.2877498600325443, a61 = -14.09640773051259, a62 = 6.925207756232704,
             a63 = -41.47510893210728, a64 = 2.343771018586405, a65 = 24.13215229196062;
const double C21 = -10.31323885133993, C31 = -21.04823117650003, C32 = -7.234992135176716, C41 = 32.22751541853323,
             C42 = -4.943732386540191, C43 = 19.44922031041879, C51 = -20.69865579590063, C52 = -8.816374604402768,
             C53 = 1.260436877740897, C54 = -0.7495647613787146, C61 = -46.22004352711257, C62 = -17.49534862857472,
             C63 = -289.6389582892057, C64 = 93.60855400400906, C65 = 318.3822534212147, C71 = 34.20013733472935,
             C72 = -14.15535402717690, C73 = 57.82335640988400, C74 = 25.83362985412365, C75 = 1.408950972071624,
             C76 = -6.551835421242162, C81 = 42.57076742291101, C82 = -13.80770672017997, C83 = 93.98938432427124,
             C84 = 18.77919633714503, C85 = -31.58359187223370, C86 = -6.685968952921985, C87 = -5.810979938412932;
const double c2 = 0.38, c3 = 0.3878509998321533, c4 = 0.4839718937873840, c5 = 0.4570477008819580;
const double d1 = gamma, d2 = -0.1823079225333714636, d3 = -0.319231832186874912, d4 = 0.3449828624725343,
             d5 = -0.377417564392089818;
}  // namespace r5


template <class T>
struct Work {
  int n;
  std::vector<T> J, LU, dT, f, u, rhs;
  std::vector<int> piv;
  std::vector<T> K[8];
  std::vector<D<T>> yd, fd;
  explicit Work(int n_) : n(n_), J(n_ * n_), LU(n_ * n_), dT(n_), f(n_), u(n_), rhs(n_), piv(n_), yd(n_), fd(n_) {
    for (auto &v : K) v.resize(n_);
  }
};

// exact J = df/dy and dT = df/dtau by forward-mode duals (jax.jacfwd)
template <class T, class P>
void jacobian(const ModelT<P> &m, const T &t, const T *y, Work<T> &w) {
  const int n = w.n;
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) w.yd[i] = D<T>(y[i], T(i == j ? 1.0 : 0.0));
    rhs<D<T>, P>(m, D<T>(t), w.yd.data(), w.fd.data());
    for (int i = 0; i < n; ++i) w.J[i * n + j] = w.fd[i].d;
  }
  for (int i = 0; i < n; ++i) w.yd[i] = D<T>(y[i]);
  rhs<D<T>, P>(m, D<T>(t, T(1.0)), w.yd.data(), w.fd.data());
  for (int i = 0; i < n; ++i) w.dT[i] = w.fd[i].d;
}

// LAPACK getrf-style partial pivoting (first max |.| in the column)
template <class T>
void lu_factor(T *A, int *piv, int n) {
  for (int kcol = 0; kcol < n; ++kcol) {
    int p = kcol;
    double best = std::fabs(val(A[kcol * n + kcol]));
    for (int i = kcol + 1; i < n; ++i) {
      double v = std::fabs(val(A[i * n + kcol]));
      if (v > best) { best = v; p = i; }
    }
    piv[kcol] = p;
    if (p != kcol)
      for (int j = 0; j < n; ++j) std::swap(A[kcol * n + j], A[p * n + j]);
    T d = A[kcol * n + kcol];
    if (val(d) != 0.0) {
      T inv = 1.0 / d;
      for (int i = kcol + 1; i < n; ++i) {
        if (val(A[i * n + kcol]) == 0.0 && val(A[i * n + kcol] * inv) == 0.0) {
          // structurally zero entry: keep the (possibly nonzero) tangent exact
          T l = A[i * n + kcol] * inv;
          A[i * n + kcol] = l;
          const T *rk = A + kcol * n;
          T *ri = A + i * n;
          for (int j = kcol + 1; j < n; ++j) ri[j] = ri[j] - l * rk[j];
          continue;
        }
        T l = A[i * n + kcol] * inv;
        A[i * n + kcol] = l;
        const T *rk = A + kcol * n;
        T *ri = A + i * n;
        for (int j = kcol + 1; j < n; ++j) ri[j] = ri[j] - l * rk[j];
      }
    }
  }
}

template <class T>
void lu_solve(const T *LU, const int *piv, int n, T *b) {
  for (int i = 0; i < n; ++i)
    if (piv[i] != i) std::swap(b[i], b[piv[i]]);
  for (int i = 1; i < n; ++i) {
    T s = b[i];
    const T *r = LU + i * n;
    for (int j = 0; j < i; ++j) s = s - r[j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    T s = b[i];
    const T *r = LU + i * n;
    for (int j = i + 1; j < n; ++j) s = s - r[j] * b[j];
    b[i] = s / r[i];
  }
}

// problem adaptor: the perturbation system of one mode
template <class P>
struct ModeSystem {
  const ModelT<P> &m;
  template <class T>
  void f(const T &t, const T *y, T *out) const { rhs<T, P>(m, t, y, out); }
  template <class T>
  void jac(const T &t, const T *y, Work<T> &w) const { jacobian<T, P>(m, t, y, w); }
};

// linear test system f = A y + g0 + g1 t (pins the step on closed forms)
struct LinearSystem {
  int n;
  const double *A, *g0, *g1;
  void f(double t, const double *y, double *out) const {
    for (int i = 0; i < n; ++i) {
      double s = g0[i] + g1[i] * t;
      for (int j = 0; j < n; ++j) s += A[i * n + j] * y[j];
      out[i] = s;
    }
  }
  void jac(double, const double *, Work<double> &w) const {
    for (int i = 0; i < n * n; ++i) w.J[i] = A[i];
    for (int i = 0; i < n; ++i) w.dT[i] = g1[i];
  }
};

// one Rodas5Transformed step; returns y1 in y1, error estimate (k8) in err
template <class T, class Sys>
void rodas5_step_sys(const Sys &sys, const T &t0, const T &t1, const T *y0, T *y1, T *err, Work<T> &w) {
  using namespace r5;
  const int n = w.n;
  const T dt = t1 - t0;
  sys.jac(t0, y0, w);
  const T dtg = dt * gamma;
  for (int i = 0; i < n * n; ++i) w.LU[i] = -w.J[i];
  for (int i = 0; i < n; ++i) w.LU[i * n + i] = w.LU[i * n + i] + 1.0 / dtg;
  lu_factor(w.LU.data(), w.piv.data(), n);
  auto &k1 = w.K[0], &k2 = w.K[1], &k3 = w.K[2], &k4 = w.K[3], &k5 = w.K[4], &k6 = w.K[5], &k7 = w.K[6], &k8 = w.K[7];
  auto F = [&](const T &t, const T *u) { sys.f(t, u, w.f.data()); };
  auto solve = [&](std::vector<T> &k) {
    std::copy(w.rhs.begin(), w.rhs.end(), k.begin());
    lu_solve(w.LU.data(), w.piv.data(), n, k.data());
  };
  F(t0, y0);
  for (int i = 0; i < n; ++i) w.rhs[i] = w.f[i] + dt * d1 * w.dT[i];
  solve(k1);
  for (int i = 0; i < n; ++i) w.u[i] = y0[i] + a21 * k1[i];
  F(t0 + c2 * dt, w.u.data());
  for (int i = 0; i < n; ++i) w.rhs[i] = w.f[i] + dt * d2 * w.dT[i] + (C21 / dt) * k1[i];
  solve(k2);
  for (int i = 0; i < n; ++i) w.u[i] = y0[i] + a31 * k1[i] + a32 * k2[i];
  F(t0 + c3 * dt, w.u.data());
  for (int i = 0; i < n; ++i) w.rhs[i] = w.f[i] + dt * d3 * w.dT[i] + ((C31 / dt) * k1[i] + (C32 / dt) * k2[i]);
  solve(k3);
  for (int i = 0; i < n; ++i) w.u[i] = y0[i] + a41 * k1[i] + a42 * k2[i] + a43 * k3[i];
  F(t0 + c4 * dt, w.u.data());
  for (int i = 0; i < n; ++i)
    w.rhs[i] = w.f[i] + dt * d4 * w.dT[i] + ((C41 / dt) * k1[i] + (C42 / dt) * k2[i] + (C43 / dt) * k3[i]);
  solve(k4);
  for (int i = 0; i < n; ++i) w.u[i] = y0[i] + a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i];
  F(t0 + c5 * dt, w.u.data());
  for (int i = 0; i < n; ++i)
    w.rhs[i] = w.f[i] + dt * d5 * w.dT[i] +
               ((C51 / dt) * k1[i] + (C52 / dt) * k2[i] + (C53 / dt) * k3[i] + (C54 / dt) * k4[i]);
  solve(k5);
  for (int i = 0; i < n; ++i) w.u[i] = y0[i] + a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i];
  F(t0 + dt, w.u.data());
  for (int i = 0; i < n; ++i)
    w.rhs[i] = w.f[i] + ((C61 / dt) * k1[i] + (C62 / dt) * k2[i] + (C63 / dt) * k3[i] + (C64 / dt) * k4[i] +
                         (C65 / dt) * k5[i]);
  solve(k6);
  for (int i = 0; i < n; ++i) w.u[i] = w.u[i] + k6[i];
  F(t0 + dt, w.u.data());
  for (int i = 0; i < n; ++i)
    w.rhs[i] = w.f[i] + ((C71 / dt) * k1[i] + (C72 / dt) * k2[i] + (C73 / dt) * k3[i] + (C74 / dt) * k4[i] +
                         (C75 / dt) * k5[i] + (C76 / dt) * k6[i]);
  solve(k7);
  for (int i = 0; i < n; ++i) w.u[i] = w.u[i] + k7[i];
  F(t0 + dt, w.u.data());
  for (int i = 0; i < n; ++i)
    w.rhs[i] = w.f[i] + ((C81 / dt) * k1[i] + (C82 / dt) * k2[i] + (C83 / dt) * k3[i] + (C84 / dt) * k4[i] +
                         (C85 / dt) * k5[i] + (C86 / dt) * k6[i] + (C87 / dt) * k7[i]);
  solve(k8);
  for (int i = 0; i < n; ++i) {
    y1[i] = w.u[i] + k8[i];
    err[i] = k8[i];
  }
}

template <class T, class P>
void rodas5_step(const ModelT<P> &m, const T &t0, const T &t1, const T *y0, T *y1, T *err, Work<T> &w) {
  rodas5_step_sys<T>(ModeSystem<P>{m}, t0, t1, y0, y1, err, w);
}

inline bool all_finite(const std::vector<double> &y) {
  for (double v : y)
    if (!std::isfinite(v)) return false;
  return true;
}

struct SolverCfg {
  double rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin;
  int max_steps;
  int return_full;
};

// Sum of squares of the filtered, weighted scaled error: rms_norm_filtered
// (perturbations.py:701-705, indices a, eta, delta_c, delta_b, theta_b,
// delta_gamma with weights 1, k^2, 1, 1, 1/k^2, 1) of diffrax's scaled error
// err / (atol + rtol max(|y0|, |y1|)).  Entries outside the filter are never read.
double filtered_sq(const double *err, const double *y0, const double *y1, double k, const SolverCfg &s) {
  const int fidx[6] = {0, 2, 3, 5, 6, 7};
  const double fw[6] = {1.0, k * k, 1.0, 1.0, 1.0 / (k * k), 1.0};
  double ss = 0.0;
  for (int j = 0; j < 6; ++j) {
    const int i = fidx[j];
    const double sc = err[i] / (s.atol + std::max(std::fabs(y0[i]), std::fabs(y1[i])) * s.rtol);
    const double x = sc * fw[j];
    ss += x * x;
  }
  return ss;
}

// diffrax PIDController.adapt_step_size for one attempted step: accept if the
// norm is < 1; a NaN error counts as inf (rejected, step shrunk by factormin);
// inv is the inverse norm carried to the next steps (reset to 1 when 0 or inf)
struct Control {
  bool keep;
  double factor, inv;
};
Control pid_control(double nrm, double prev_inv, double prev_prev_inv, const SolverCfg &s) {
  const double eo = 5.0;
  const double coeff1 = (s.icoeff + s.pcoeff + s.dcoeff) / eo;
  const double coeff2 = -(s.pcoeff + 2.0 * s.dcoeff) / eo;
  const double coeff3 = s.dcoeff / eo;
  if (!std::isfinite(nrm)) nrm = INFINITY;
  Control c;
  c.keep = nrm < 1.0;
  double inv = 1.0 / nrm;
  const double f1 = coeff1 == 0.0 ? 1.0 : std::pow(inv, coeff1);
  const double f2 = coeff2 == 0.0 ? 1.0 : std::pow(prev_inv, coeff2);
  const double f3 = coeff3 == 0.0 ? 1.0 : std::pow(prev_prev_inv, coeff3);
  const double fmin = c.keep ? 1.0 : s.factormin;
  c.factor = std::min(std::max(0.9 * f1 * f2 * f3, fmin), s.factormax);
  if (inv == 0.0 || std::isinf(inv)) inv = 1.0;
  c.inv = inv;
  return c;
}

// evolve_one_mode: returns status; fills out[nout * (return_full ? n : 20)].
// times (optional) receives t0 and the end time of every accepted step.
int evolve_mode(const Model &m, const SolverCfg &s, const double *tau_out, int nout, double *out, int *nsteps,
                int *naccept, double *tau_start_out, std::vector<double> *times = nullptr) {
  const int n = m.nvar();
  Work<double> w(n);
  double tmin = tau_out[0];
  for (int i = 1; i < nout; ++i) tmin = std::min(tmin, tau_out[i]);
  double tmax = tau_out[0];
  for (int i = 1; i < nout; ++i) tmax = std::max(tmax, tau_out[i]);
  const double t0 = start_time<double>(m, tmin);
  if (tau_start_out) *tau_start_out = t0;
  if (times) times->assign(1, t0);
  const double t1 = tmax;
  std::vector<double> y(n), y1(n), err(n), ys(n);
  ics<double>(m, t0, y.data());
  const double dt0 = std::min(t0 / 4.0, 0.5 * (t1 - t0));
  double tprev = t0, tnext = std::min(t0 + dt0, t1);
  double prev_inv = 1.0, prev_prev_inv = 1.0;
  int iout = 0, steps = 0, acc = 0;
  const int width = s.return_full ? n : 20;
  int status = DEB_STATUS_OK;
  while (tprev < t1) {
    if (steps >= s.max_steps) { status = DEB_STATUS_MAX_STEPS; break; }
    rodas5_step<double, double>(m, tprev, tnext, y.data(), y1.data(), err.data(), w);
    const double nrm = std::sqrt(filtered_sq(err.data(), y.data(), y1.data(), m.k, s) / 6.0);
    const Control ctl = pid_control(nrm, prev_inv, prev_prev_inv, s);
    const bool keep = ctl.keep;
    const double inv = ctl.inv;
    double dt = (tnext - tprev) * ctl.factor;
    if (keep) {
      while (iout < nout && tau_out[iout] <= tnext) {
        double th = (tau_out[iout] - tprev) / (tnext - tprev);
        for (int i = 0; i < n; ++i) ys[i] = y[i] + th * (y1[i] - y[i]);
        if (s.return_full) std::copy(ys.begin(), ys.end(), out + (size_t)iout * width);
        else outputs<double, double>(m, ys.data(), out + (size_t)iout * width);
        ++iout;
      }
      prev_prev_inv = prev_inv;
      prev_inv = inv;
      y.swap(y1);
      ++acc;
      if (times) times->push_back(tnext);
    }
    double nt0 = keep ? tnext : tprev;
    double nt1 = nt0 + dt;
    if (nt1 > t1 - 1e-10) nt1 = keep ? t1 : nt0 + 0.5 * (t1 - nt0);
    tprev = nt0;
    tnext = nt1;
    ++steps;
    // a non-finite accepted state, or a step size shrunk to nothing
    if (keep ? !all_finite(y) : !(nt1 > nt0)) { status = DEB_STATUS_NONFINITE; break; }
  }
  // outputs a failed mode never reached hold inf (diffrax's unreached SaveAt entries)
  if (status != DEB_STATUS_OK) std::fill(out + (size_t)iout * width, out + (size_t)nout * width, INFINITY);
  *nsteps = steps;
  *naccept = acc;
  return status;
}

// evolve_modes_batched for one group of modes (perturbations.py:820-922):
// the group's start time is the minimum of its modes' start times
// (determine_start_times_per_batch :786-803), every mode starts from its ICs
// there, and one step sequence serves the group: the controller sees the RMS
// over all modes' filtered, weighted errors (rms_norm_filtered_batched
// :690-694).  out[b][nout][width]; every mode gets the group's status.
int evolve_batch(const std::vector<Model> &ms, const SolverCfg &s, const double *tau_out, int nout, double *out,
                 int *nsteps, int *naccept, std::vector<double> *trace = nullptr) {
  const int B = (int)ms.size();
  const int n = ms[0].nvar();
  Work<double> w(n);
  double tmin = tau_out[0], tmax = tau_out[0];
  for (int i = 1; i < nout; ++i) {
    tmin = std::min(tmin, tau_out[i]);
    tmax = std::max(tmax, tau_out[i]);
  }
  double t0 = start_time<double>(ms[0], tmin);
  for (int b = 1; b < B; ++b) t0 = std::min(t0, start_time<double>(ms[b], tmin));
  const double t1 = tmax;
  std::vector<std::vector<double>> y(B, std::vector<double>(n)), y1 = y;
  std::vector<double> err(n), ys(n);
  for (int b = 0; b < B; ++b) ics<double>(ms[b], t0, y[b].data());
  const double dt0 = std::min(t0 / 4.0, 0.5 * (t1 - t0));
  double tprev = t0, tnext = std::min(t0 + dt0, t1);
  double prev_inv = 1.0, prev_prev_inv = 1.0;
  int iout = 0, steps = 0, acc = 0;
  const int width = s.return_full ? n : 20;
  int status = DEB_STATUS_OK;
  while (tprev < t1) {
    if (steps >= s.max_steps) { status = DEB_STATUS_MAX_STEPS; break; }
    // rms_norm_filtered_batched (:690-694): one RMS over every mode's 6 entries
    double ss = 0.0;
    for (int b = 0; b < B; ++b) {
      rodas5_step<double, double>(ms[b], tprev, tnext, y[b].data(), y1[b].data(), err.data(), w);
      ss += filtered_sq(err.data(), y[b].data(), y1[b].data(), ms[b].k, s);
    }
    const double nrm = std::sqrt(ss / (6.0 * B));
    if (trace) trace->insert(trace->end(), {tprev, tnext, nrm});
    const Control ctl = pid_control(nrm, prev_inv, prev_prev_inv, s);
    const bool keep = ctl.keep;
    const double inv = ctl.inv;
    double dt = (tnext - tprev) * ctl.factor;
    if (keep) {
      while (iout < nout && tau_out[iout] <= tnext) {
        double th = (tau_out[iout] - tprev) / (tnext - tprev);
        for (int b = 0; b < B; ++b) {
          for (int i = 0; i < n; ++i) ys[i] = y[b][i] + th * (y1[b][i] - y[b][i]);
          double *o = out + ((size_t)b * nout + iout) * width;
          if (s.return_full) std::copy(ys.begin(), ys.end(), o);
          else outputs<double, double>(ms[b], ys.data(), o);
        }
        ++iout;
      }
      prev_prev_inv = prev_inv;
      prev_inv = inv;
      for (int b = 0; b < B; ++b) y[b].swap(y1[b]);
      ++acc;
    }
    double nt0 = keep ? tnext : tprev;
    double nt1 = nt0 + dt;
    if (nt1 > t1 - 1e-10) nt1 = keep ? t1 : nt0 + 0.5 * (t1 - nt0);
    tprev = nt0;
    tnext = nt1;
    ++steps;
    bool bad = !keep && !(nt1 > nt0);
    for (int b = 0; keep && b < B; ++b) bad = bad || !all_finite(y[b]);
    if (bad) { status = DEB_STATUS_NONFINITE; break; }
  }
  if (status != DEB_STATUS_OK)
    for (int b = 0; b < B; ++b)
      std::fill(out + ((size_t)b * nout + iout) * width, out + ((size_t)b * nout + nout) * width, INFINITY);
  *nsteps = steps;
  *naccept = acc;
  return status;
}

// directional derivative of the 20 output fields along one tangent of the
// inputs, replaying the accepted step end times `times` (t0 first) with each
// step held fixed as a fraction of [t0, t1] (see the file header)
void replay_tangent(const ModelT<Dd> &md, const std::vector<double> &times, const double *tau_out,
                    const double *dtau_out, int nout, double *out, double *dout) {
  const int n = md.nvar();
  std::vector<Dd> tout(nout);
  for (int i = 0; i < nout; ++i) tout[i] = Dd(tau_out[i], dtau_out[i]);
  Dd tmin = tout[0], tmax = tout[0];
  for (int i = 1; i < nout; ++i) {
    if (tau_out[i] < val(tmin)) tmin = tout[i];
    if (tau_out[i] > val(tmax)) tmax = tout[i];
  }
  const Dd T0 = start_time<Dd>(md, tmin);
  const Dd T1 = tmax;
  const double span = times.back() - times.front();
  auto time_of = [&](size_t i) {
    double sfr = (times[i] - times.front()) / span;
    return Dd(times[i], (1.0 - sfr) * T0.d + sfr * T1.d);
  };
  Work<Dd> w(n);
  std::vector<Dd> y(n), y1(n), err(n), ys(n), o(20);
  ics<Dd>(md, time_of(0), y.data());
  int iout = 0;
  for (size_t st_i = 0; st_i + 1 < times.size(); ++st_i) {
    const Dd ta = time_of(st_i), tb = time_of(st_i + 1);
    rodas5_step<Dd, Dd>(md, ta, tb, y.data(), y1.data(), err.data(), w);
    while (iout < nout && tau_out[iout] <= times[st_i + 1]) {
      Dd th = (tout[iout] - ta) / (tb - ta);
      for (int i = 0; i < n; ++i) ys[i] = y[i] + th * (y1[i] - y[i]);
      outputs<Dd, Dd>(md, ys.data(), o.data());
      for (int f = 0; f < 20; ++f) {
        out[(size_t)iout * 20 + f] = o[f].v;
        dout[(size_t)iout * 20 + f] = o[f].d;
      }
      ++iout;
    }
    y.swap(y1);
  }
}

int evolve_tangent(const Model &m, const ModelT<Dd> &md, const SolverCfg &s, const double *tau_out,
                   const double *dtau_out, int nout, double *out, double *dout, int *naccept) {
  std::vector<double> times, plain((size_t)nout * 20);
  int nsteps = 0, nacc = 0;
  double ts0 = 0.0;
  SolverCfg s20 = s;
  s20.return_full = 0;
  int st = evolve_mode(m, s20, tau_out, nout, plain.data(), &nsteps, &nacc, &ts0, &times);
  *naccept = nacc;
  if (st != DEB_STATUS_OK) return st;
  replay_tangent(md, times, tau_out, dtau_out, nout, out, dout);
  return DEB_STATUS_OK;
}

template <class P>
void model_common(ModelT<P> &m, const deb_cosmo_t *c, const deb_config_t *cfg, const double *q, const double *w,
                  double k) {
  m.k = k;
  m.lmaxg = cfg->lmaxg;
  m.lmaxgp = cfg->lmaxgp;
  m.lmaxr = cfg->lmaxr;
  m.lmaxnu = cfg->lmaxnu;
  m.nq = cfg->nqmax;
  for (int i = 0; i < m.nq; ++i) {
    m.q[i] = q[i];
    m.w[i] = w[i];
  }
  m.lambda = (c->w_DE_0 == -1.0 && c->w_DE_a == 0.0);
}

Model make_model(const deb_cosmo_t *c, const double *tables, const int32_t *desc, const deb_config_t *cfg,
                 const double *q, const double *w, double k) {
  Model m;
  m.c = cosmo_of(*c);
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    int off = desc[2 * s], n = desc[2 * s + 1];
    m.sp[s] = {tables + off, tables + off + n, tables + off + 2 * n, n};
  }
  model_common(m, c, cfg, q, w, k);
  return m;
}

// tangent-carrying model; knots (value, tangent) live in `store`
ModelT<Dd> make_model_tangent(const deb_cosmo_t *c, const double *tables, const deb_cosmo_t *dc,
                              const double *dtables, const int32_t *desc, const deb_config_t *cfg, const double *q,
                              const double *w, double k, std::vector<Dd> &store) {
  ModelT<Dd> m;
  m.c = cosmo_of(*c, *dc);
  size_t tot = 0;
  for (int s = 0; s < DEB_NSPLINE; ++s) tot += 3 * (size_t)desc[2 * s + 1];
  store.resize(tot);
  size_t pos = 0;
  for (int s = 0; s < DEB_NSPLINE; ++s) {
    int off = desc[2 * s], n = desc[2 * s + 1];
    for (int j = 0; j < 3 * n; ++j) store[pos + j] = Dd(tables[off + j], dtables[off + j]);
    m.sp[s] = {store.data() + pos, store.data() + pos + n, store.data() + pos + 2 * n, n};
    pos += 3 * (size_t)n;
  }
  model_common(m, c, cfg, q, w, k);
  return m;
}

}  // namespace

extern "C" {

// Rodas5 W-form tableau: a[8][8] (row i: stage-i combination of k_j; row 8
// also gives y1 = u8 + k8), C[8][8], c[8], d[8], gamma
int oracle_rodas5_tableau(double *a, double *C, double *c, double *d, double *gamma) {
  using namespace r5;
  std::fill(a, a + 64, 0.0);
  std::fill(C, C + 64, 0.0);
  const double A[8][8] = {{0}, {a21}, {a31, a32}, {a41, a42, a43}, {a51, a52, a53, a54},
                          {a61, a62, a63, a64, a65}, {a61, a62, a63, a64, a65, 1.0},
                          {a61, a62, a63, a64, a65, 1.0, 1.0}};
  const double CC[8][8] = {{0}, {C21}, {C31, C32}, {C41, C42, C43}, {C51, C52, C53, C54},
                           {C61, C62, C63, C64, C65}, {C71, C72, C73, C74, C75, C76},
                           {C81, C82, C83, C84, C85, C86, C87}};
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) {
      a[i * 8 + j] = A[i][j];
      C[i * 8 + j] = CC[i][j];
    }
  const double cc[8] = {0.0, c2, c3, c4, c5, 1.0, 1.0, 1.0};
  const double dd[8] = {d1, d2, d3, d4, d5, 0.0, 0.0, 0.0};
  for (int i = 0; i < 8; ++i) {
    c[i] = cc[i];
    d[i] = dd[i];
  }
  *gamma = r5::gamma;
  return 0;
}

int oracle_nvar(const deb_config_t *cfg) {
  return 12 + cfg->lmaxg + cfg->lmaxgp + cfg->lmaxr + cfg->nqmax * (cfg->lmaxnu + 1);
}

// f = model_synchronous(tau, y) for one mode
int oracle_rhs(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
               const double *q, const double *w, double k, double tau, const double *y, double *f) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  rhs<double, double>(m, tau, y, f);
  return 0;
}

// exact dense Jacobian J[n*n] (row-major) and dT = df/dtau
int oracle_jacobian(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                    const double *q, const double *w, double k, double tau, const double *y, double *J, double *dT) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  Work<double> wk(m.nvar());
  jacobian<double, double>(m, tau, y, wk);
  std::copy(wk.J.begin(), wk.J.end(), J);
  std::copy(wk.dT.begin(), wk.dT.end(), dT);
  return 0;
}

int oracle_rodas5_step(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                       const double *q, const double *w, double k, double t0, double t1, const double *y0, double *y1,
                       double *err) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  Work<double> wk(m.nvar());
  rodas5_step<double, double>(m, t0, t1, y0, y1, err, wk);
  return 0;
}

int oracle_ics(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
               const double *q, const double *w, double k, double tau_out_min, double *tau_start, double *y0) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  *tau_start = start_time<double>(m, tau_out_min);
  ics<double>(m, *tau_start, y0);
  return 0;
}

// one Rodas5 step of the linear system y' = A y + g0 + g1 t (same step code)
int oracle_rodas5_linear(int n, const double *A, const double *g0, const double *g1, double t0, double t1,
                         const double *y0, double *y1, double *err) {
  Work<double> wk(n);
  rodas5_step_sys<double>(LinearSystem{n, A, g0, g1}, t0, t1, y0, y1, err, wk);
  return 0;
}

// The step-size controller's view of one attempted step: out[0] = error norm
// (filtered RMS), out[1] = 1 if accepted, out[2] = step-size factor, out[3] =
// the inverse norm carried forward.
int oracle_step_control(const deb_config_t *cfg, double k, const double *err, const double *y0, const double *y1,
                        double prev_inv, double prev_prev_inv, double *out) {
  SolverCfg s{cfg->rtol, cfg->atol, cfg->pcoeff, cfg->icoeff, cfg->dcoeff, cfg->factormax, cfg->factormin,
              cfg->max_steps, cfg->return_full};
  const double nrm = std::sqrt(filtered_sq(err, y0, y1, k, s) / 6.0);
  const Control c = pid_control(nrm, prev_inv, prev_prev_inv, s);
  out[0] = std::isfinite(nrm) ? nrm : INFINITY;
  out[1] = c.keep ? 1.0 : 0.0;
  out[2] = c.factor;
  out[3] = c.inv;
  return 0;
}

// The first nk modes as one group of evolve_modes_batched; trace[3 i + 0..2]
// = (t0, t1, error norm) of attempted step i for i < max_trace; returns the
// number of attempted steps (negative on error).
int oracle_batch_trace(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                       const double *q, const double *w, int nk, const double *kmodes, int nout,
                       const double *tau_out, int max_trace, double *trace) {
  SolverCfg s{cfg->rtol, cfg->atol, cfg->pcoeff, cfg->icoeff, cfg->dcoeff, cfg->factormax, cfg->factormin,
              cfg->max_steps, cfg->return_full};
  std::vector<Model> ms;
  for (int b = 0; b < nk; ++b) ms.push_back(make_model(c, tables, desc, cfg, q, w, kmodes[b]));
  const int n = oracle_nvar(cfg);
  std::vector<double> out((size_t)nk * nout * (cfg->return_full ? n : 20));
  std::vector<double> tr;
  int ns = 0, na = 0;
  evolve_batch(ms, s, tau_out, nout, out.data(), &ns, &na, &tr);
  const int m = std::min<int>(max_trace, (int)tr.size() / 3);
  std::copy(tr.begin(), tr.begin() + 3 * m, trace);
  return ns;
}

// adiabatic initial conditions at an explicit time tau
int oracle_ics_at(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                  const double *q, const double *w, double k, double tau, double *y0) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  ics<double>(m, tau, y0);
  return 0;
}

int oracle_outputs(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                   const double *q, const double *w, double k, const double *y, double *out20) {
  Model m = make_model(c, tables, desc, cfg, q, w, k);
  outputs<double, double>(m, y, out20);
  return 0;
}

// Full per-mode solve for nk modes of one cosmology, threaded over modes.
int oracle_evolve(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                  const double *q, const double *w, int nk, const double *kmodes, int nout, const double *tau_out,
                  double *y_out, int32_t *status, int32_t *nsteps, int32_t *naccept, double *tau_start,
                  int nthreads, double *seconds) {
  SolverCfg s{cfg->rtol, cfg->atol, cfg->pcoeff, cfg->icoeff, cfg->dcoeff, cfg->factormax, cfg->factormin,
              cfg->max_steps, cfg->return_full};
  const int n = oracle_nvar(cfg);
  const size_t width = cfg->return_full ? (size_t)n : 20;
  std::atomic<int> next(0);
  auto worker = [&]() {
    for (;;) {
      int i = next.fetch_add(1);
      if (i >= nk) break;
      const auto clk = std::chrono::steady_clock::now();
      Model m = make_model(c, tables, desc, cfg, q, w, kmodes[i]);
      int st, na;
      double ts;
      status[i] = evolve_mode(m, s, tau_out, nout, y_out + (size_t)i * nout * width, &st, &na, &ts);
      if (seconds) seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - clk).count();
      nsteps[i] = st;
      naccept[i] = na;
      if (tau_start) tau_start[i] = ts;
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
  }
  return 0;
}

// evolve_modes_batched: groups of `batch` consecutive k (nk % batch == 0)
int oracle_evolve_batched(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables, const int32_t *desc,
                          const double *q, const double *w, int nk, const double *kmodes, int nout,
                          const double *tau_out, int batch, double *y_out, int32_t *status, int32_t *nsteps,
                          int32_t *naccept, int nthreads) {
  if (batch < 1 || nk % batch != 0) return -1;
  SolverCfg s{cfg->rtol, cfg->atol, cfg->pcoeff, cfg->icoeff, cfg->dcoeff, cfg->factormax, cfg->factormin,
              cfg->max_steps, cfg->return_full};
  const int n = oracle_nvar(cfg);
  const size_t width = cfg->return_full ? (size_t)n : 20;
  const int ng = nk / batch;
  std::atomic<int> next(0);
  auto worker = [&]() {
    for (;;) {
      int g = next.fetch_add(1);
      if (g >= ng) break;
      std::vector<Model> ms;
      for (int b = 0; b < batch; ++b) ms.push_back(make_model(c, tables, desc, cfg, q, w, kmodes[g * batch + b]));
      int st, na;
      const int r = evolve_batch(ms, s, tau_out, nout, y_out + (size_t)g * batch * nout * width, &st, &na);
      for (int b = 0; b < batch; ++b) {
        status[g * batch + b] = r;
        nsteps[g * batch + b] = st;
        naccept[g * batch + b] = na;
      }
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
  }
  return 0;
}

// ---- tangents (directional derivatives along dcosmo, dtables, dtau, dy) ----

// f and its tangent
int oracle_rhs_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                       const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                       const double *w, double k, double tau, double dtau, const double *y, const double *dy,
                       double *f, double *df) {
  std::vector<Dd> store;
  ModelT<Dd> m = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, k, store);
  const int n = m.nvar();
  std::vector<Dd> yd(n), fd(n);
  for (int i = 0; i < n; ++i) yd[i] = Dd(y[i], dy[i]);
  rhs<Dd, Dd>(m, Dd(tau, dtau), yd.data(), fd.data());
  for (int i = 0; i < n; ++i) {
    f[i] = fd[i].v;
    df[i] = fd[i].d;
  }
  return 0;
}

// one Rodas5 step and its tangent along (dt0, dt1, dy0, dcosmo, dtables)
int oracle_rodas5_step_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                               const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                               const double *w, double k, double t0, double dt0, double t1, double dt1,
                               const double *y0, const double *dy0, double *y1, double *dy1) {
  std::vector<Dd> store;
  ModelT<Dd> m = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, k, store);
  const int n = m.nvar();
  Work<Dd> wk(n);
  std::vector<Dd> a(n), b(n), e(n);
  for (int i = 0; i < n; ++i) a[i] = Dd(y0[i], dy0[i]);
  rodas5_step<Dd, Dd>(m, Dd(t0, dt0), Dd(t1, dt1), a.data(), b.data(), e.data(), wk);
  for (int i = 0; i < n; ++i) {
    y1[i] = b[i].v;
    dy1[i] = b[i].d;
  }
  return 0;
}

// 20 output fields and their tangent along (dy, dcosmo, dtables)
int oracle_outputs_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                           const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                           const double *w, double k, const double *y, const double *dy, double *out20,
                           double *dout20) {
  std::vector<Dd> store;
  ModelT<Dd> m = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, k, store);
  const int n = m.nvar();
  std::vector<Dd> yd(n), o(20);
  for (int i = 0; i < n; ++i) yd[i] = Dd(y[i], dy[i]);
  outputs<Dd, Dd>(m, yd.data(), o.data());
  for (int f = 0; f < 20; ++f) {
    out20[f] = o[f].v;
    dout20[f] = o[f].d;
  }
  return 0;
}

// start time, initial state and their tangents along (dtau_out_min, dcosmo, dtables)
int oracle_ics_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                       const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                       const double *w, double k, double tau_out_min, double dtau_out_min, double *tau_start,
                       double *dtau_start, double *y0, double *dy0) {
  std::vector<Dd> store;
  ModelT<Dd> m = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, k, store);
  const int n = m.nvar();
  Dd ts = start_time<Dd>(m, Dd(tau_out_min, dtau_out_min));
  std::vector<Dd> y(n);
  ics<Dd>(m, ts, y.data());
  *tau_start = ts.v;
  *dtau_start = ts.d;
  for (int i = 0; i < n; ++i) {
    y0[i] = y[i].v;
    dy0[i] = y[i].d;
  }
  return 0;
}

// Per-mode solve plus the tangent of its 20 output fields along one input
// tangent (dcosmo, dtables, dtau_out), threaded over modes.
int oracle_evolve_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                          const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                          const double *w, int nk, const double *kmodes, int nout, const double *tau_out,
                          const double *dtau_out, double *y_out, double *dy_out, int32_t *status, int32_t *naccept,
                          int nthreads, double *seconds) {
  SolverCfg s{cfg->rtol, cfg->atol, cfg->pcoeff, cfg->icoeff, cfg->dcoeff, cfg->factormax, cfg->factormin,
              cfg->max_steps, 0};
  std::atomic<int> next(0);
  auto worker = [&]() {
    std::vector<Dd> store;
    for (;;) {
      int i = next.fetch_add(1);
      if (i >= nk) break;
      const auto clk = std::chrono::steady_clock::now();
      Model m = make_model(c, tables, desc, cfg, q, w, kmodes[i]);
      ModelT<Dd> md = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, kmodes[i], store);
      int na = 0;
      status[i] = evolve_tangent(m, md, s, tau_out, dtau_out, nout, y_out + (size_t)i * nout * 20,
                                 dy_out + (size_t)i * nout * 20, &na);
      if (seconds) seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - clk).count();
      naccept[i] = na;
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
  }
  return 0;
}

// tangent of the outputs along one input tangent for a given sequence of
// accepted step end times (times[0] = start time); see replay_tangent
int oracle_replay_tangent(const deb_config_t *cfg, const deb_cosmo_t *c, const double *tables,
                          const deb_cosmo_t *dc, const double *dtables, const int32_t *desc, const double *q,
                          const double *w, double k, int ntimes, const double *times, int nout,
                          const double *tau_out, const double *dtau_out, double *y_out, double *dy_out) {
  std::vector<Dd> store;
  ModelT<Dd> md = make_model_tangent(c, tables, dc, dtables, desc, cfg, q, w, k, store);
  std::vector<double> tv(times, times + ntimes);
  replay_tangent(md, tv, tau_out, dtau_out, nout, y_out, dy_out);
  return 0;
}

}  // extern "C"
