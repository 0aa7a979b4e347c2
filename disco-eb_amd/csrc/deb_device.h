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

// pointers into a wave's LDS (address space 3: ds_read/ds_write, also inside
// outlined functions, which would otherwise address LDS with flat instructions)
#ifndef DEB_LDS
#define DEB_LDS __attribute__((address_space(3)))
#endif
typedef DEB_LDS double lds_double;

// doubles of LDS one wave needs for its step factors: the cyclic-reduction
// coefficients (alpha, beta) of every level and row
template <int NS>
constexpr int pcr_lds_doubles() {
  return PCR_LEVELS * 2 * NS * WAVE;
}

// the forward kernels keep the step's 7 stage vectors in LDS where the wave's
// share of it allows (four waves per CU: 40 KB each, with the PCR
// coefficients), else in registers; the same LDS stages the output fields
// between steps
template <int NS>
constexpr bool stage_in_lds() {
  return (7 * NS * WAVE + pcr_lds_doubles<NS>()) * 8 <= 40960;
}
template <int NS>
constexpr int fwd_lds_doubles() {
  return stage_in_lds<NS>() ? 7 * NS * WAVE : NS * WAVE;
}

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
__device__ __forceinline__ dual &operator+=(dual &a, dual b) {
  a.v += b.v;
  a.d += b.d;
  return a;
}
__device__ __forceinline__ dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator-(dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator/(dual a, dual b) {
  const double ib = 1.0 / b.v;
  const double q = a.v * ib;
  return {q, (a.d - q * b.d) * ib};
}
__device__ __forceinline__ dual operator+(dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ dual operator+(double b, dual a) { return {a.v + b, a.d}; }
__device__ __forceinline__ dual operator-(dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ dual operator-(double b, dual a) { return {b - a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ dual operator*(double b, dual a) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ dual operator/(double b, dual a) {
  const double ia = 1.0 / a.v;
  const double q = b * ia;
  return {q, -q * a.d * ia};
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

// 1/x and 1/sqrt(x) for positive normal x: the hardware estimate refined by
// two Newton steps, within about one ulp (the IEEE division and square root
// add scaling and a final correction step that these per-lane row constants
// do not need)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * r, r, 0.5);
  r = fma(r, e, r);
  e = fma(-h * r, r, 0.5);
  return fma(r, e, r);
}
__device__ __forceinline__ double vval(double a) { return a; }
__device__ __forceinline__ double vval(dual a) { return a.v; }
__device__ __forceinline__ double dval(double) { return 0.0; }
__device__ __forceinline__ double dval(dual a) { return a.d; }

// ----------------------------------------------------------------------------
// multi-tangent dual number: value v and N directional derivatives d[0..N-1].
// B is double, or `dual` for a second-order (nested) derivative.
template <int N, class B = double>
struct dn {
  static constexpr int size = N;
  B v;
  B d[N];
  __device__ __forceinline__ dn &operator=(double x) {
    v = x;
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
    return *this;
  }
};
template <int N, class B>
__device__ __forceinline__ dn<N, B> mkdn(B v) {
  dn<N, B> r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator+(const dn<N, B> &a, const dn<N, B> &b) {
  dn<N, B> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator-(const dn<N, B> &a, const dn<N, B> &b) {
  dn<N, B> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator-(const dn<N, B> &a) {
  dn<N, B> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator*(const dn<N, B> &a, const dn<N, B> &b) {
  dn<N, B> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator/(const dn<N, B> &a, const dn<N, B> &b) {
  dn<N, B> r;
  B ib = 1.0 / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator+(const dn<N, B> &a, double b) {
  dn<N, B> r = a;
  r.v = a.v + b;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator+(double b, const dn<N, B> &a) { return a + b; }
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator-(const dn<N, B> &a, double b) {
  dn<N, B> r = a;
  r.v = a.v - b;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator-(double b, const dn<N, B> &a) {
  dn<N, B> r;
  r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator*(const dn<N, B> &a, double b) {
  dn<N, B> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator*(double b, const dn<N, B> &a) { return a * b; }
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator/(const dn<N, B> &a, double b) { return a * (1.0 / b); }
template <int N, class B>
__device__ __forceinline__ dn<N, B> operator/(double b, const dn<N, B> &a) {
  dn<N, B> r;
  B ia = 1.0 / a.v;
  r.v = b * ia;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -r.v * a.d[i] * ia;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> dexp(const dn<N, B> &a) {
  dn<N, B> r;
  r.v = dexp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> dlog(const dn<N, B> &a) {
  dn<N, B> r;
  r.v = dlog(a.v);
  B ia = 1.0 / a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ia;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> dsqrt(const dn<N, B> &a) {
  dn<N, B> r;
  r.v = dsqrt(a.v);
  B h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}
template <int N, class B>
__device__ __forceinline__ dn<N, B> dpow(const dn<N, B> &a, double p) {
  dn<N, B> r;
  r.v = dpow(a.v, p);
  B g = p * dpow(a.v, p - 1.0);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = g * a.d[i];
  return r;
}
template <int N, class B>
__device__ __forceinline__ double vval(const dn<N, B> &a) { return vval(a.v); }

template <class T>
__device__ __forceinline__ T tconst(double v) {
  T r;
  r = v;
  return r;
}

template <class T>
struct is_dn { static constexpr bool value = false; };
template <int N, class B>
struct is_dn<dn<N, B>> { static constexpr bool value = true; };
// scalar type of the cosmology parameters and spline knots when computing in
// T: tangent-carrying for dn, plain double otherwise
template <class T>
struct par_of { using type = double; };
template <int N, class B>
struct par_of<dn<N, B>> { using type = dn<N, B>; };
template <class T>
using ParT = typename par_of<T>::type;

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
// value of an arbitrary lane (byte address = 4 * lane)
__device__ __forceinline__ double from_lane(double x, int src) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_ds_bpermute(src << 2, lo);
  hi = __builtin_amdgcn_ds_bpermute(src << 2, hi);
  return __hiloint2double(hi, lo);
}
#ifdef DEB_NO_DPP
__device__ __forceinline__ double from_prev(double x) {
  const int l = __lane_id();
  double v = from_lane(x, l > 0 ? l - 1 : 0);
  return l > 0 ? v : 0.0;
}
__device__ __forceinline__ double from_next(double x) {
  const int l = __lane_id();
  double v = from_lane(x, l < 63 ? l + 1 : 63);
  return l < 63 ? v : 0.0;
}
#else
// value of lane i-1 (lane 0 receives 0)
__device__ __forceinline__ double from_prev(double x) { return dpp_f64<0x138, false>(x); }
// value of lane i+1 (lane 63 receives 0)
__device__ __forceinline__ double from_next(double x) { return dpp_f64<0x130, false>(x); }
#endif

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

// Wave-wide sums by a symmetric butterfly: xor 1 and xor 2 (quad_perm), the
// half-row and row mirrors, then xor 16 and xor 32 (ds_bpermute).  Every
// partial sum pairs two equal-sized groups, so each lane ends with the same
// bits; the results stay in vector registers (no scalar-register pressure).
#ifdef DEB_NO_DPP
__device__ __forceinline__ double row_sum16(double x) {
  const int l = __lane_id();
  x += from_lane(x, l ^ 1);
  x += from_lane(x, l ^ 2);
  x += from_lane(x, l ^ 4);
  x += from_lane(x, l ^ 8);
  return x;
}
#else
__device__ __forceinline__ double row_sum16(double x) {
  x += dpp_f64<0xB1, true>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E, true>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141, true>(x);  // row_half_mirror
  x += dpp_f64<0x140, true>(x);  // row_mirror
  return x;
}
#endif
template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N]) {
  const int lane = lane_id();
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = row_sum16(v[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += from_lane(v[i], lane ^ 16);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += from_lane(v[i], lane ^ 32);
}
// one wave-wide sum, returned as a wave-uniform (scalar) value
__device__ __forceinline__ double wave_sum1(double x) {
  double v[1] = {x};
  wave_sum<1>(v);
  return readfirst(v[0]);
}

// ----------------------------------------------------------------------------
// checkpoint tape of the gradient path.  The accepted states n = 0..N of a mode
// are grouped in segments of C consecutive steps: segment s keeps the
// checkpoint state y_{sC} and the times t_{sC} .. t_{sC+C}.  One chunk per
// segment, handed out from one pool by an atomic counter and linked backwards
// (a mode's segments are read back last-to-first).  The reverse launch
// recomputes each segment forward from its checkpoint.
//   chunk:  [int64 prev chunk (-1: none)][pad][double t[C + 1]][pad to 256 B][double y[NS][64]]
//   buffer: [header 256 B: uint64 chunks handed out][int64 last[n_modes]][int32 count[n_modes]][pool]
constexpr int TAPE_MAX_CKPT = 4096;

__host__ __device__ inline long long tape_state_off(int ckpt) { return (64 + 8LL * (ckpt + 1) + 255) / 256 * 256; }
__host__ __device__ inline long long tape_chunk_bytes(int ns, int ckpt) {
  return tape_state_off(ckpt) + 8LL * ns * WAVE;
}

struct TapeView {
  unsigned long long *counter;  // chunks handed out so far
  long long capacity;           // chunks in the pool
  long long *last;              // [n_modes] chunk of a mode's last segment
  int32_t *count;               // [n_modes] states recorded (-1: pool exhausted)
  char *pool;
  long long chunk_bytes;
  long long state_off;          // byte offset of the checkpoint state in a chunk
  int ckpt;                     // steps per segment C
};

// byte layout of a tape buffer for n_modes modes: header, per-mode arrays, pool
__host__ __device__ inline long long tape_last_off() { return 256; }
__host__ __device__ inline long long tape_count_off(long long n_modes) { return 256 + 8 * n_modes; }
__host__ __device__ inline long long tape_pool_off(long long n_modes) {
  long long b = 256 + 12 * n_modes;
  return (b + 255) / 256 * 256;
}

// Per-wave ring of the reverse launch: the states y_{sC} .. y_{sC+L} and the
// eight stage vectors of every step of the segment being reversed, each a
// lane-contiguous [NS][64] block.
__host__ __device__ inline long long ring_doubles(int ns, int ckpt) {
  return (long long)ns * WAVE * (9LL * ckpt + 1);
}

// ----------------------------------------------------------------------------
// natural cubic spline lookups on the per-cosmology tables (uniform per wave)
struct Spl {
  const double *x, *y, *S;
  int n;
  int off;  // offset of the knots in the cosmology's table block
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

// cubic spline piece i at xv (knots x, y, S of s)
template <typename T>
__device__ __forceinline__ T spl_piece(const Spl &s, int i, T xv) {
  double x0 = s.x[i], x1 = s.x[i + 1];
  double h = x1 - x0;
  T t = (xv - x0) / h;
  T A = 1.0 - t;
  T B = t;
  double y0 = s.y[i], y1 = s.y[i + 1], s0 = s.S[i], s1 = s.S[i + 1];
  return A * y0 + B * y1 + ((A * A * A - A) * s0 + (B * B * B - B) * s1) * (h * h / 6.0);
}

// a wave-uniform pointer moved to scalar registers, so that loads through it
// with a wave-uniform index can go through the scalar cache
__device__ __forceinline__ const double *uniform_ptr(const double *p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return (const double *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// cubic spline piece from preloaded knots (x0, x1, y0, y1, S0, S1)
template <typename T>
__device__ __forceinline__ T spl_piece_k(const double (&kn)[6], T xv) {
  const double h = kn[1] - kn[0];
  T t = (xv - kn[0]) / h;
  T A = 1.0 - t;
  T B = t;
  return A * kn[2] + B * kn[3] + ((A * A * A - A) * kn[4] + (B * B * B - B) * kn[5]) * (h * h / 6.0);
}

// three splines at one wave-uniform abscissa (the same intervals as
// spl_interval): all six knots of the three hinted intervals are loaded
// before any is tested (one round trip when the hints hold, which is the rule
// between the stages of a step); a missed hint searches and reloads
template <typename T>
__device__ __forceinline__ void spl_eval3(const Spl (&sp)[3], T xv, int *const (&hint)[3], T (&r)[3]) {
  const double v = vval(xv);
  int i[3];
  double kn[3][6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    i[j] = *hint[j];
    int c = i[j] < 0 ? 0 : (i[j] > sp[j].n - 2 ? sp[j].n - 2 : i[j]);
    c = __builtin_amdgcn_readfirstlane(c);
    const double *x = uniform_ptr(sp[j].x), *y = uniform_ptr(sp[j].y), *S = uniform_ptr(sp[j].S);
    kn[j][0] = x[c];
    kn[j][1] = x[c + 1];
    kn[j][2] = y[c];
    kn[j][3] = y[c + 1];
    kn[j][4] = S[c];
    kn[j][5] = S[c + 1];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const bool ok = i[j] >= 0 && i[j] <= sp[j].n - 2 && kn[j][0] < v && v <= kn[j][1];
    if (!ok) {
      const int c = spl_interval(sp[j], v, *hint[j]);
      kn[j][0] = sp[j].x[c];
      kn[j][1] = sp[j].x[c + 1];
      kn[j][2] = sp[j].y[c];
      kn[j][3] = sp[j].y[c + 1];
      kn[j][4] = sp[j].S[c];
      kn[j][5] = sp[j].S[c + 1];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) r[j] = spl_piece_k(kn[j], xv);
}

template <typename T>
__device__ __forceinline__ T spl_eval(const Spl &s, T xv, int &hint) {
  return spl_piece(s, spl_interval(s, vval(xv), hint), xv);
}

}  // namespace deb
