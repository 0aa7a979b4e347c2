// Device building blocks of the gfx950 perturbation kernels.
//
// One wavefront integrates one k-mode.  The state vector is spread over the 64
// lanes as NS "slots" (one double register per slot per lane); every
// hierarchy is a contiguous run of lanes inside one slot, so neighbouring
// multipoles l-1, l+1 sit on lanes i-1, i+1 of the same register and the
// tridiagonal part of W = I/(gamma h) - J is solved by parallel cyclic
// reduction across lanes.
//
// Slot 0, lanes 0..7 hold a, a*h' accumulator, eta, delta_gamma, delta_c,
// theta_c, delta_Q, theta_Q; lanes 8.. hold the photon-temperature chain
// (delta_b, theta_b, theta_gamma, F_2 .. F_lmaxg).  The other chains
// (polarization G_0..G_lmaxgp, massless nu, one chain per massive-nu momentum
// bin) are packed by the host into the remaining lanes/slots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/discoeb_hip.h"

namespace deb {

constexpr int WAVE = 64;
constexpr int MAXNS = 6;
constexpr int PCR_LEVELS = 6;

// row types (lane roles)
enum : uint32_t {
  R_NONE = 0,
  R_A, R_AHP, R_ETA, R_DG, R_DC, R_TC, R_DQ, R_TQ,
  R_DB, R_TB, R_TG, R_F2, R_FL, R_FLMAX,
  R_G0, R_GL, R_G2, R_GLMAX,
  R_DR, R_TR, R_FR2, R_FRL, R_FRLMAX,
  R_Q0, R_Q1, R_Q2, R_QL, R_QLMAX,
  R_NTYPES
};

// fixed lanes of slot 0
constexpr int L_A = 0, L_AHP = 1, L_ETA = 2, L_DG = 3, L_DC = 4, L_TC = 5, L_DQ = 6, L_TQ = 7, L_DB = 8, L_TB = 9,
              L_TG = 10;

// row descriptor bits
__host__ __device__ inline uint32_t rd_type(uint32_t d) { return d & 31u; }
__host__ __device__ inline uint32_t rd_l(uint32_t d) { return (d >> 5) & 127u; }
__host__ __device__ inline uint32_t rd_q(uint32_t d) { return (d >> 12) & 15u; }
__host__ __device__ inline uint32_t rd_idx(uint32_t d) { return (d >> 16) & 1023u; }
__host__ __device__ inline uint32_t rd_make(uint32_t t, uint32_t l, uint32_t q, uint32_t idx) {
  return t | (l << 5) | (q << 12) | (idx << 16);
}

// kernel parameter block (passed by value)
template <int NS>
struct Params {
  deb_config_t cfg;
  int n, nq;
  int n_cosmo, nk, nout;
  int64_t table_stride;
  int32_t desc[DEB_NSPLINE][2];
  double q[DEB_MAXNQ], w[DEB_MAXNQ], dlnf[DEB_MAXNQ];
  const deb_cosmo_t *cosmo;
  const double *tables;
  const double *kmodes;
  const double *tau_out;
  double *y_out;
  int32_t *status;
  int32_t *nsteps;
  uint32_t rowdesc[NS][WAVE];
};

// ----------------------------------------------------------------------------
// forward-mode dual number (derivative with respect to the scale factor)
struct dual {
  double v, d;
  __device__ __forceinline__ dual &operator=(double x) {
    v = x;
    d = 0.0;
    return *this;
  }
};
__device__ __forceinline__ dual mkd(double v, double d = 0.0) { return {v, d}; }
__device__ __forceinline__ dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator-(dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator/(dual a, dual b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ dual operator+(dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ dual operator+(double b, dual a) { return {a.v + b, a.d}; }
__device__ __forceinline__ dual operator-(dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ dual operator-(double b, dual a) { return {b - a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ dual operator*(double b, dual a) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ dual operator/(double b, dual a) {
  double q = b / a.v;
  return {q, -q * a.d / a.v};
}
__device__ __forceinline__ dual dexp(dual a) {
  double e = exp(a.v);
  return {e, e * a.d};
}
__device__ __forceinline__ dual dlog(dual a) { return {log(a.v), a.d / a.v}; }
__device__ __forceinline__ dual dsqrt(dual a) {
  double s = sqrt(a.v);
  return {s, 0.5 * a.d / s};
}
__device__ __forceinline__ dual dpow(dual a, double p) {
  double v = pow(a.v, p);
  return {v, p * v / a.v * a.d};
}
__device__ __forceinline__ double dexp(double a) { return exp(a); }
__device__ __forceinline__ double dlog(double a) { return log(a); }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double dpow(double a, double p) { return pow(a, p); }
__device__ __forceinline__ double vval(double a) { return a; }
__device__ __forceinline__ double vval(dual a) { return a.v; }
__device__ __forceinline__ double dval(double) { return 0.0; }
__device__ __forceinline__ double dval(dual a) { return a.d; }

// ----------------------------------------------------------------------------
// cross-lane primitives (doubles moved as two 32-bit halves)
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <int CTRL, bool BOUND_ZERO>
__device__ __forceinline__ double dpp_f64(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, BOUND_ZERO);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, BOUND_ZERO);
  return __hiloint2double(hi, lo);
}
// value of lane i-1 (lane 0 receives 0)
__device__ __forceinline__ double from_prev(double x) { return dpp_f64<0x138, false>(x); }
// value of lane i+1 (lane 63 receives 0)
__device__ __forceinline__ double from_next(double x) { return dpp_f64<0x130, false>(x); }

// value of an arbitrary lane (byte address = 4 * lane)
__device__ __forceinline__ double from_lane(double x, int src) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_ds_bpermute(src << 2, lo);
  hi = __builtin_amdgcn_ds_bpermute(src << 2, hi);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane(double x, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readfirst(double x) {
  int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
  int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// sum over the 16 lanes of each DPP row, result in every lane of the row
__device__ __forceinline__ double row_sum16(double x) {
  x += dpp_f64<0xB1, true>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E, true>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x124, true>(x);  // row_ror:4
  x += dpp_f64<0x128, true>(x);  // row_ror:8
  return x;
}
// wave-uniform sum of N values
template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = row_sum16(v[i]);
#pragma unroll
  for (int i = 0; i < N; ++i)
    v[i] = (readlane(v[i], 0) + readlane(v[i], 16)) + (readlane(v[i], 32) + readlane(v[i], 48));
}
__device__ __forceinline__ double wave_sum1(double x) {
  double v[1] = {x};
  wave_sum<1>(v);
  return v[0];
}

// ----------------------------------------------------------------------------
// natural cubic spline lookups on the per-cosmology tables (uniform per wave)
struct Spl {
  const double *x, *y, *S;
  int n;
};

// interval i with x[i] < v <= x[i+1], clipped to [0, n-2]
// (clip(searchsorted(x, v, 'left') - 1, 0, n - 2)); `hint` caches the last interval
__device__ __forceinline__ int spl_interval(const Spl &s, double v, int &hint) {
  int i = hint;
  if (i >= 0 && i <= s.n - 2) {
    if (s.x[i] < v && v <= s.x[i + 1]) return i;
    if (i + 2 <= s.n - 1 && s.x[i + 1] < v && v <= s.x[i + 2]) {
      hint = i + 1;
      return i + 1;
    }
    if (i >= 1 && s.x[i - 1] < v && v <= s.x[i]) {
      hint = i - 1;
      return i - 1;
    }
  }
  int lo = 0, hi = s.n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (s.x[mid] < v) lo = mid + 1; else hi = mid;
  }
  i = lo - 1;
  i = i < 0 ? 0 : i;
  i = i > s.n - 2 ? s.n - 2 : i;
  hint = i;
  return i;
}

template <typename T>
__device__ __forceinline__ T spl_eval(const Spl &s, T xv, int &hint) {
  int i = spl_interval(s, vval(xv), hint);
  double x0 = s.x[i], x1 = s.x[i + 1];
  double h = x1 - x0;
  T t = (xv - x0) / h;
  T A = 1.0 - t;
  T B = t;
  double y0 = s.y[i], y1 = s.y[i + 1], s0 = s.S[i], s1 = s.S[i + 1];
  return A * y0 + B * y1 + ((A * A * A - A) * s0 + (B * B * B - B) * s1) * (h * h / 6.0);
}

}  // namespace deb
