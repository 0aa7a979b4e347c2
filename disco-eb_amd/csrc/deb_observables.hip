// Observables of the perturbation outputs on the GPU: the power spectrum
// (get_power, perturbations.py:1101-1123), its Savitzky-Golay smoothed version
// (get_power_smoothed :1126-1159 with util.py:407-444 savgol_filter), the
// Kaiser redshift-space spectrum (power_Kaiser :1162-1198), its multipoles
// (power_multipoles :1202-1224) and the FFTlog correlation function
// (get_xi_from_P :1065-1098 with util.py:12-44 lngamma_complex_e).
//
// One workgroup per cosmology; a cosmology's spectrum (nk <= DEB_OBS_MAXK)
// lives in LDS while it is filtered or transformed.  All FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "deb_host.h"

namespace deb {

constexpr int OBS_BLOCK = 256;
constexpr double PI = 3.14159265358979323846;

// 2 pi^2 A_s (k/k_p)^(n_s - 1) k^-3: the primordial factor of get_power
__device__ __forceinline__ double prim(double k, double A_s, double n_s, double k_p) {
  return 2.0 * PI * PI * A_s * pow(k / k_p, n_s - 1.0) / (k * k * k);
}

__global__ void __launch_bounds__(OBS_BLOCK)
    deb_power_kernel(int n_cosmo, int nk, const double *k, const double *y, int nfield, int idx, const double *A_s,
                     const double *n_s, const double *k_p, double *P) {
  const int64_t i = (int64_t)blockIdx.x * OBS_BLOCK + threadIdx.x;
  if (i >= (int64_t)n_cosmo * nk) return;
  const int c = (int)(i / nk);
  const double v = y[i * nfield + idx];
  P[i] = prim(k[i], A_s[c], n_s[c], k_p[c]) * v * v;
}

// Savitzky-Golay smoothing coefficients (savgol_filter, util.py:407-444): the
// minimum-norm solution c of A c = e_0 with A[m][j] = x_j^m, x_j = pos - j
// (the reference's reversed grid), i.e. c = Q R^-T e_0 for the thin QR
// factorisation A^T = Q R.  Modified Gram-Schmidt with one reorthogonalisation
// pass on x scaled to [-1, 1] (the scaling of row m cancels in c).  One
// thread; p + 1 <= 8 columns of length W.
__device__ void savgol_coeffs(int W, int p, double *c, double *q /* [p+1][W] */) {
  const int half = W / 2;
  const double pos = (W % 2) ? (double)half : half - 0.5;
  const double sc = pos > 0.0 ? 1.0 / pos : 1.0;
  double R[8][8];
  for (int m = 0; m <= p; ++m) {
    double *qm = q + m * W;
    for (int j = 0; j < W; ++j) qm[j] = pow((pos - j) * sc, (double)m);
    for (int n = 0; n <= m; ++n) R[n][m] = 0.0;
    for (int pass = 0; pass < 2; ++pass)
      for (int n = 0; n < m; ++n) {
        const double *qn = q + n * W;
        double d = 0.0;
        for (int j = 0; j < W; ++j) d += qn[j] * qm[j];
        for (int j = 0; j < W; ++j) qm[j] -= d * qn[j];
        R[n][m] += d;
      }
    double nn = 0.0;
    for (int j = 0; j < W; ++j) nn += qm[j] * qm[j];
    nn = sqrt(nn);
    R[m][m] = nn;
    for (int j = 0; j < W; ++j) qm[j] /= nn;
  }
  // z = R^-T e_0 (forward substitution with R^T, lower triangular)
  double z[8];
  for (int m = 0; m <= p; ++m) {
    double s = (m == 0) ? 1.0 : 0.0;
    for (int n = 0; n < m; ++n) s -= R[n][m] * z[n];
    z[m] = s / R[m][m];
  }
  for (int j = 0; j < W; ++j) {
    double s = 0.0;
    for (int m = 0; m <= p; ++m) s += q[m * W + j] * z[m];
    c[j] = s;
  }
}

// exp(convolve(log P, c, 'same')) with the first W//2 and the last
// ceil(W/2) entries replaced by P (get_power_smoothed :1153-1157); sp holds
// log P on entry (LDS), out receives the result
__device__ void smooth_log(int nk, int W, const double *c, const double *lp, double *out, const double *Pm) {
  const int h = (W - 1) / 2;
  const int head = W / 2, tail = (W + 1) / 2;
  for (int i = threadIdx.x; i < nk; i += blockDim.x) {
    double s;
    if (i < head || i >= nk - tail) {
      s = Pm[i];
    } else {
      double acc = 0.0;
      // full convolution index n = i + h: sum_j c[j] lp[n - j]
      for (int j = 0; j < W; ++j) {
        const int t = i + h - j;
        if (t >= 0 && t < nk) acc += c[j] * lp[t];
      }
      s = exp(acc);
    }
    out[i] = s;
  }
}

// LDS of one cosmology: P, log P, the coefficients and the QR scratch
template <int MAXK>
struct ObsLds {
  double P[MAXK], lp[MAXK], c[MAXK];
};

__global__ void __launch_bounds__(OBS_BLOCK)
    deb_power_smoothed_kernel(int nk, const double *k, const double *y, int nfield, int idx, const double *A_s,
                              const double *n_s, const double *k_p, int W, int p, double *q_scratch, double *out) {
  __shared__ ObsLds<DEB_OBS_MAXK> s;
  const int c = blockIdx.x;
  const double *kc = k + (int64_t)c * nk;
  for (int i = threadIdx.x; i < nk; i += blockDim.x) {
    const double v = y[((int64_t)c * nk + i) * nfield + idx];
    const double P = prim(kc[i], A_s[c], n_s[c], k_p[c]) * v * v;
    s.P[i] = P;
    s.lp[i] = log(P);
  }
  if (threadIdx.x == 0) savgol_coeffs(W, p, s.c, q_scratch + (int64_t)c * (p + 1) * W);
  __syncthreads();
  smooth_log(nk, W, s.c, s.lp, out + (int64_t)c * nk, s.P);
}

// (b delta - mu^2 theta)^2 (power_Kaiser :1195-1197).  W = 0: delta, theta
// from the unsmoothed fields 4 and 5 times sqrt of the primordial factor;
// W > 0: delta = sqrt(smoothed P_delta), theta = -sqrt(smoothed P_theta)
__global__ void __launch_bounds__(OBS_BLOCK)
    deb_power_kaiser_kernel(int nk, int nmu, const double *k, const double *y, int nfield, const double *A_s,
                            const double *n_s, const double *k_p, const double *bias, const double *mu, int W, int p,
                            double *q_scratch, double *out) {
  __shared__ ObsLds<DEB_OBS_MAXK> s;
  __shared__ double dl[DEB_OBS_MAXK], th[DEB_OBS_MAXK];
  const int c = blockIdx.x;
  const double *kc = k + (int64_t)c * nk;
  const double *yc = y + (int64_t)c * nk * nfield;
  if (W == 0) {
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const double f = sqrt(prim(kc[i], A_s[c], n_s[c], k_p[c]));
      dl[i] = f * yc[(int64_t)i * nfield + 4];
      th[i] = f * yc[(int64_t)i * nfield + 5];
    }
  } else {
    if (threadIdx.x == 0) savgol_coeffs(W, p, s.c, q_scratch + (int64_t)c * (p + 1) * W);
    for (int pass = 0; pass < 2; ++pass) {
      const int idx = 4 + pass;
      double *dst = pass == 0 ? dl : th;
      __syncthreads();
      for (int i = threadIdx.x; i < nk; i += blockDim.x) {
        const double v = yc[(int64_t)i * nfield + idx];
        const double P = prim(kc[i], A_s[c], n_s[c], k_p[c]) * v * v;
        s.P[i] = P;
        s.lp[i] = log(P);
      }
      __syncthreads();
      smooth_log(nk, W, s.c, s.lp, dst, s.P);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      dl[i] = sqrt(dl[i]);
      th[i] = -sqrt(th[i]);
    }
  }
  __syncthreads();
  const double b = bias[c];
  for (int t = threadIdx.x; t < nk * nmu; t += blockDim.x) {
    const int i = t / nmu, j = t % nmu;
    const double v = b * dl[i] - mu[j] * mu[j] * th[i];
    out[((int64_t)c * nk + i) * nmu + j] = v * v;
  }
}

// P0, P2, P4 of the Kaiser spectrum (power_multipoles :1212-1224)
__global__ void __launch_bounds__(OBS_BLOCK)
    deb_power_multipoles_kernel(int n_cosmo, int nk, const double *k, const double *y, int nfield,
                                const double *A_s, const double *n_s, const double *k_p, const double *bias,
                                double *P0, double *P2, double *P4) {
  const int64_t i = (int64_t)blockIdx.x * OBS_BLOCK + threadIdx.x;
  if (i >= (int64_t)n_cosmo * nk) return;
  const int c = (int)(i / nk);
  const double f = sqrt(prim(k[i], A_s[c], n_s[c], k_p[c]));
  const double d = f * y[i * nfield + 4], t = f * y[i * nfield + 5], b = bias[c];
  P0[i] = b * b * d * d - 2.0 * b / 3.0 * d * t + 1.0 / 5.0 * t * t;
  P2[i] = -4.0 * b / 3.0 * d * t + 4.0 / 7.0 * t * t;
  P4[i] = 8.0 / 35.0 * t * t;
}

// ---------------------------------------------------------------------------
// FFTlog (get_xi_from_P)

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx clog(cplx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }

// Im log Gamma(z) for Re z > 0.5 by the Lanczos series of lngamma_complex_e
// (util.py:12-44, after GSL): with z' = z - 1,
// (z' + 1/2) log(z' + 7.5) - (z' + 7.5) + log sqrt(2 pi) + log A_g(z'),
// A_g = c_0 + sum_{i=1..8} c_i conj(z' + i) / |z' + i|^2; for Re z <= 0.5 the
// reflection log pi - log sin(pi z) - lnGamma(1 - z)
__device__ cplx lngamma_lanczos(cplx z) {
  const double cl[9] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                        -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                        -176.61502916214059906584551354,     12.507343278686904814458936853,
                        -0.13857109526572011689554707,       9.984369578019570859563e-6,
                        1.50563273514931155834e-7};
  z.re -= 1.0;
  cplx ag = {cl[0], 0.0};
  for (int i = 1; i <= 8; ++i) {
    const double xr = z.re + i, xi = z.im;
    const double d = xr * xr + xi * xi;
    ag.re += cl[i] * xr / d;
    ag.im -= cl[i] * xi / d;
  }
  const cplx l75 = clog({z.re + 7.5, z.im});
  const cplx t = cmul({z.re + 0.5, z.im}, l75);
  const cplx la = clog(ag);
  return {t.re - (z.re + 7.5) + 0.9189385332046727418 + la.re, t.im - z.im + la.im};
}

__device__ cplx lngamma_complex(cplx z) {
  if (z.re > 0.5) return lngamma_lanczos(z);
  // sin(pi z) = sin(pi x) cosh(pi y) + i cos(pi x) sinh(pi y)
  const double x = PI * z.re, yv = PI * z.im;
  const cplx s = {sin(x) * cosh(yv), cos(x) * sinh(yv)};
  const cplx ls = clog(s);
  const cplx lg = lngamma_lanczos({1.0 - z.re, -z.im});
  return {1.14472988584940017414342735135 - ls.re - lg.re, -ls.im - lg.im};
}

// cos and sin of 2 pi m / N for integer m (reduced mod N first, so the
// argument is exact to the rounding of one division)
__device__ __forceinline__ cplx twiddle(int64_t m, int N) {
  m %= N;
  if (m < 0) m += N;
  double s, c;
  sincos(2.0 * PI * (double)m / (double)N, &s, &c);
  return {c, s};
}

// per cosmology: f_m = rfft(P k^1.5)[m] (m = 0..N/2) by direct summation,
// f_m *= exp(2i (theta_m - log(pi) kappa_m)), kappa_m = pi m / L,
// theta_m = Im lnGamma((1.5 + ell)/2 + i kappa_m); then
// xi_n = Re(i^ell irfft(f)_n) / (2 pi r_n)^1.5 with r = 2 pi / k, returned
// in reversed order (r increasing) with r
__global__ void __launch_bounds__(OBS_BLOCK)
    deb_xi_kernel(int N, const double *k, const double *P, int ell, double *xi, double *r) {
  __shared__ double a[DEB_OBS_MAXK];
  __shared__ cplx f[DEB_OBS_MAXK / 2 + 1];
  const int c = blockIdx.x;
  const double *kc = k + (int64_t)c * N;
  const double *Pc = P + (int64_t)c * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = Pc[i] * pow(kc[i], 1.5);
  __syncthreads();
  const double L = log(kc[N - 1] / kc[0]);
  const int M = N / 2 + 1;
  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < N; ++j) {
      const cplx w = twiddle(-(int64_t)m * j, N);
      re += a[j] * w.re;
      im += a[j] * w.im;
    }
    const double kap = PI * m / L;
    const double theta = lngamma_complex({(1.5 + ell) / 2.0, kap}).im;
    double s, co;
    sincos(2.0 * (theta - log(PI) * kap), &s, &co);
    f[m] = cmul({re, im}, {co, s});
  }
  __syncthreads();
  // irfft of length N: x_n = (1/N) [f_0 + sum_{m=1}^{N/2-1} 2 Re(f_m e^{2 pi i m n / N}) + Re(f_{N/2} e^{i pi n})]
  // (numpy drops the imaginary parts of f_0 and of f_{N/2})
  const int e4 = ((ell % 4) + 4) % 4;  // i^ell = 1, i, -1, -i
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    double x = f[0].re;
    for (int m = 1; m < N / 2; ++m) {
      const cplx w = twiddle((int64_t)m * n, N);
      x += 2.0 * (f[m].re * w.re - f[m].im * w.im);
    }
    x += f[N / 2].re * ((n & 1) ? -1.0 : 1.0);
    x /= N;
    // i^ell times a real number: its real part is x for ell = 0 mod 4, -x for
    // 2 mod 4, and 0 for odd ell
    const double re = e4 == 0 ? x : (e4 == 2 ? -x : 0.0);
    const double rn = 2.0 * PI / kc[n];
    const double v = re / pow(2.0 * PI * rn, 1.5);
    xi[(int64_t)c * N + (N - 1 - n)] = v;
    r[(int64_t)c * N + (N - 1 - n)] = rn;
  }
}

// the QR scratch of the smoothing coefficients: (polyorder + 1) x W doubles per cosmology
static size_t deb_obs_scratch_bytes_impl(int n_cosmo, int W, int p) {
  if (n_cosmo < 1 || W < 1 || p < 0) return 0;
  return (size_t)n_cosmo * (size_t)(p + 1) * (size_t)W * sizeof(double);
}

static int check_common(int n_cosmo, int nk, const void *k, const void *out) {
  if (n_cosmo < 1 || nk < 1) return fail(DEB_E_ARG, "n_cosmo and nk must be positive");
  if (!k || !out) return fail(DEB_E_ARG, "null pointer");
  return 0;
}

static int check_params(const double *A_s, const double *n_s, const double *k_p, const double *y, int nfield) {
  if (!A_s || !n_s || !k_p || !y) return fail(DEB_E_ARG, "null pointer");
  if (nfield < 1) return fail(DEB_E_ARG, "nfield must be positive");
  return 0;
}

static int check_window(int nk, int W, int p, size_t scratch_bytes, int n_cosmo) {
  if (nk > DEB_OBS_MAXK) return fail(DEB_E_UNSUPPORTED, "nk above DEB_OBS_MAXK");
  if (W < 1 || W > nk) return fail(DEB_E_ARG, "window_length outside [1, nk]");
  if (p < 0 || p > 7 || p >= W) return fail(DEB_E_ARG, "polyorder outside [0, min(7, window_length - 1)]");
  if (scratch_bytes < deb_obs_scratch_bytes_impl(n_cosmo, W, p))
    return fail(DEB_E_ARG, "scratch smaller than deb_observables_scratch_bytes");
  return 0;
}

}  // namespace deb

using namespace deb;

extern "C" {

size_t deb_observables_scratch_bytes(int n_cosmo, int window_length, int polyorder) {
  return deb_obs_scratch_bytes_impl(n_cosmo, window_length, polyorder);
}

int deb_power(int n_cosmo, int nk, const double *k, const double *y, int nfield, int idx, const double *A_s,
              const double *n_s, const double *k_p, double *P, void *stream) {
  if (int r = check_common(n_cosmo, nk, k, P)) return r;
  if (int r = check_params(A_s, n_s, k_p, y, nfield)) return r;
  if (idx < 0 || idx >= nfield) return fail(DEB_E_ARG, "field index outside [0, nfield)");
  const int64_t n = (int64_t)n_cosmo * nk;
  hipLaunchKernelGGL(deb_power_kernel, dim3((unsigned)((n + OBS_BLOCK - 1) / OBS_BLOCK)), dim3(OBS_BLOCK), 0,
                     (hipStream_t)stream, n_cosmo, nk, k, y, nfield, idx, A_s, n_s, k_p, P);
  return hip_check(hipGetLastError(), "power kernel launch");
}

int deb_power_smoothed(int n_cosmo, int nk, const double *k, const double *y, int nfield, int idx,
                       const double *A_s, const double *n_s, const double *k_p, int window_length, int polyorder,
                       void *scratch, size_t scratch_bytes, double *P_smooth, void *stream) {
  if (int r = check_common(n_cosmo, nk, k, P_smooth)) return r;
  if (int r = check_params(A_s, n_s, k_p, y, nfield)) return r;
  if (idx < 0 || idx >= nfield) return fail(DEB_E_ARG, "field index outside [0, nfield)");
  if (int r = check_window(nk, window_length, polyorder, scratch ? scratch_bytes : 0, n_cosmo)) return r;
  hipLaunchKernelGGL(deb_power_smoothed_kernel, dim3(n_cosmo), dim3(OBS_BLOCK), 0, (hipStream_t)stream, nk, k, y,
                     nfield, idx, A_s, n_s, k_p, window_length, polyorder, (double *)scratch, P_smooth);
  return hip_check(hipGetLastError(), "smoothed power kernel launch");
}

int deb_power_kaiser(int n_cosmo, int nk, int nmu, const double *k, const double *y, int nfield,
                     const double *A_s, const double *n_s, const double *k_p, const double *bias, const double *mu,
                     int window_length, int polyorder, void *scratch, size_t scratch_bytes, double *Pkmu,
                     void *stream) {
  if (int r = check_common(n_cosmo, nk, k, Pkmu)) return r;
  if (int r = check_params(A_s, n_s, k_p, y, nfield)) return r;
  if (nfield < 6) return fail(DEB_E_ARG, "the Kaiser spectrum needs fields 4 and 5 (nfield >= 6)");
  if (nmu < 1 || !mu || !bias) return fail(DEB_E_ARG, "nmu must be positive, mu and bias non-null");
  if (nk > DEB_OBS_MAXK) return fail(DEB_E_UNSUPPORTED, "nk above DEB_OBS_MAXK");
  if (window_length != 0)
    if (int r = check_window(nk, window_length, polyorder, scratch ? scratch_bytes : 0, n_cosmo)) return r;
  hipLaunchKernelGGL(deb_power_kaiser_kernel, dim3(n_cosmo), dim3(OBS_BLOCK), 0, (hipStream_t)stream, nk, nmu, k, y,
                     nfield, A_s, n_s, k_p, bias, mu, window_length, polyorder, (double *)scratch, Pkmu);
  return hip_check(hipGetLastError(), "Kaiser power kernel launch");
}

int deb_power_multipoles(int n_cosmo, int nk, const double *k, const double *y, int nfield, const double *A_s,
                         const double *n_s, const double *k_p, const double *bias, double *P0, double *P2,
                         double *P4, void *stream) {
  if (int r = check_common(n_cosmo, nk, k, P0)) return r;
  if (int r = check_params(A_s, n_s, k_p, y, nfield)) return r;
  if (nfield < 6) return fail(DEB_E_ARG, "the multipoles need fields 4 and 5 (nfield >= 6)");
  if (!bias || !P2 || !P4) return fail(DEB_E_ARG, "null pointer");
  const int64_t n = (int64_t)n_cosmo * nk;
  hipLaunchKernelGGL(deb_power_multipoles_kernel, dim3((unsigned)((n + OBS_BLOCK - 1) / OBS_BLOCK)), dim3(OBS_BLOCK),
                     0, (hipStream_t)stream, n_cosmo, nk, k, y, nfield, A_s, n_s, k_p, bias, P0, P2, P4);
  return hip_check(hipGetLastError(), "multipoles kernel launch");
}

int deb_xi_from_power(int n_cosmo, int N, const double *k, const double *Pk, int ell, double *xi, double *r,
                      void *stream) {
  if (int rc = check_common(n_cosmo, N, k, xi)) return rc;
  if (!Pk || !r) return fail(DEB_E_ARG, "null pointer");
  if (N < 2 || N % 2) return fail(DEB_E_ARG, "FFTlog needs an even number of k points");
  if (N > DEB_OBS_MAXK) return fail(DEB_E_UNSUPPORTED, "N above DEB_OBS_MAXK");
  if (ell < 0) return fail(DEB_E_ARG, "ell must be >= 0");
  hipLaunchKernelGGL(deb_xi_kernel, dim3(n_cosmo), dim3(OBS_BLOCK), 0, (hipStream_t)stream, N, k, Pk, ell, xi, r);
  return hip_check(hipGetLastError(), "FFTlog kernel launch");
}

}  // extern "C"
