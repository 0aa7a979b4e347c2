// gfx950 kernels and C ABI of the gradient path.
//
//   deb_adjoint_kernel       one wavefront per (cosmology, k) item: reads the
//                            mode's states back last-to-first, reverses every
//                            step, the output interpolation, the initial
//                            conditions and the start time, and returns the
//                            derivative of the caller's scalar along each
//                            input tangent direction
//   deb_rhs_vjp_kernel,
//   deb_step_vjp_kernel      test hooks: reverse of one right-hand side
//                            evaluation and of one Rodas5 step
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "deb_adjoint.h"
#include "deb_host.h"

namespace deb {

// One workgroup (one wavefront) per (cosmology, k) item, largest k first.
// Each wave first takes a free ring (the state and stage-vector buffer of the
// segment it reverses) by a bounded atomic search from its own index: the host
// provides twice as many rings as waves can be resident, so the search ends
// within a few probes; a wave that finds none after 4 passes gives NaN
// gradients instead of waiting.  The ring is released when the item is done.
template <int NS>
__global__ void __launch_bounds__(WAVE) deb_adjoint_kernel(const Params<NS> P, const AdjArgs A) {
  // the stage cotangents of one reverse step; between steps the first two
  // [NS][64] blocks stage the output conversion's reverse (output_vjp)
  __shared__ double KBl[7 * NS * WAVE];
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  __shared__ Ctx cxs;
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * P.nk) return;
  int ki, ci;
  if (A.order) {  // modes by accepted steps, longest first
    const int mode = A.order[item];
    ci = mode / P.nk;
    ki = mode % P.nk;
  } else {
    ki = P.nk - 1 - item / P.n_cosmo;
    ci = item % P.n_cosmo;
  }
  const int lane = lane_id();
  int ring = -1;
  if (lane == 0) {
    const int start = item % A.n_rings;
    for (int t = 0; t < 4 * A.n_rings; ++t) {
      const int r = (start + t) % A.n_rings;
      if (atomicCAS(&A.ring_busy[r], 0, 1) == 0) {
        ring = r;
        break;
      }
    }
  }
  ring = __builtin_amdgcn_readfirstlane(ring);
  if (ring < 0) {
    if (lane < A.ndir) A.grad[((long long)ci * P.nk + ki) * A.ndir + lane] = __builtin_nan("");
    return;
  }
  adjoint_mode<NS>(P, A, ci, ki, A.rings + (long long)ring * ring_doubles(NS, A.tape.ckpt), KBl, KBl,
                   KBl + NS * WAVE, (lds_double *)pcr, cxs);
  __threadfence();
  if (lane == 0) atomicExch(&A.ring_busy[ring], 0);
}

// marks every ring free (a kernel rather than a memset, so that a caller may
// capture the launch sequence in a graph)
__global__ void deb_ring_reset_kernel(int32_t *busy, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) busy[i] = 0;
}

struct VjpHook {
  const int32_t *ci;
  const double *kk, *t0, *t1, *y, *wbar;
  double *ybar, *t0bar, *t1bar, *grad;
  int m, ndir;
  const double *dcos, *dtab;
};

template <int NS>
__device__ __forceinline__ void hook_ctx(Ctx &cx, const Params<NS> &P, const VjpHook &H, int item) {
  const int ci = H.ci[item];
  make_ctx<NS>(cx, P, ci, H.kk[item]);
  cx.ndir = H.ndir;
  cx.dcos = H.dcos + (long long)ci * H.ndir * COSMO_DOUBLES;
  cx.dtab = H.dtab + (long long)ci * H.ndir * P.table_stride;
}

template <int NS>
__global__ void __launch_bounds__(WAVE) deb_rhs_vjp_kernel(const Params<NS> P, const VjpHook H) {
  const int item = blockIdx.x;
  if (item >= H.m) return;
  const int lane = lane_id();
  Ctx cx;
  hook_ctx<NS>(cx, P, H, item);
  uint32_t rd[NS];
  double y[NS], w[NS], yb[NS], tw[NS];
  const double *ys = H.y + (long long)item * P.n, *ws = H.wbar + (long long)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rd[s] = P.rowdesc[s][lane];
    const bool on = rd_type(rd[s]) != R_NONE;
    y[s] = on ? ys[rd_idx(rd[s])] : 0.0;
    w[s] = on ? ws[rd_idx(rd[s])] : 0.0;
  }
  const double tau = H.t0[item], it = 1.0 / tau;
  const double a = readlane(y[0], L_A);
  double fv[1][FT_N], fb[1][FT_N], Yq[1][NS], wq[1][NS];
  RowSums rs[1];
  rs[0] = coef_pass<NS, true>(cx, rd, a, it, w, y, yb, fv[0]);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Yq[0][s] = y[s];
    wq[0][s] = w[s];
  }
#pragma unroll
  for (int q = 0; q < FT_N; ++q) fb[0][q] = 0.0;
  feature_partials<NS, 1>(cx, rd, fv, Yq, wq, rs, fb);
  wave_sum<FT_N>(fb[0]);
  const double gl = lane_contract(cx, a, fb[0]);
  double gth[MAXDIR];
  lane_gather(gl, H.ndir, gth);
  const double ga = readlane(gl, 0);
  if (lane == L_A) yb[0] += ga;
  trunc_weights<NS>(rd, tw);
  double pt = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) pt += w[s] * tw[s] * it * it * y[s];
  pt = wave_sum1(pt);
  double *yo = H.ybar + (long long)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    if (rd_type(rd[s]) != R_NONE) yo[rd_idx(rd[s])] = yb[s];
  if (lane == 0) H.t0bar[item] = pt;
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j)
    if (lane == j && j < H.ndir) H.grad[(long long)item * H.ndir + j] = gth[j];
}

template <int NS>
__global__ void __launch_bounds__(WAVE) deb_step_vjp_kernel(const Params<NS> P, const VjpHook H) {
  __shared__ double Kl[8 * NS * WAVE];
  __shared__ double KBl[7 * NS * WAVE];
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  const int item = blockIdx.x;
  if (item >= H.m) return;
  const int lane = lane_id();
  Ctx cx;
  hook_ctx<NS>(cx, P, H, item);
  __shared__ ModeU mu;
  Mode<NS> m;
  m.pcr = (lds_double *)pcr;
  m.u = (lds_mode_u *)&mu;
  m.nlev = mode_levels(P.cfg, P.nq);
  double y[NS], yb[NS];
  const double *ys = H.y + (long long)item * P.n, *ws = H.wbar + (long long)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    m.rd[s] = P.rowdesc[s][lane];
    const bool on = rd_type(m.rd[s]) != R_NONE;
    y[s] = on ? ys[rd_idx(m.rd[s])] : 0.0;
    yb[s] = on ? ws[rd_idx(m.rd[s])] : 0.0;
  }
  const double t0 = H.t0[item], h = H.t1[item] - t0;
  double thb[MAXDIR];
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j) thb[j] = 0.0;
  double tb = 0.0, hb = 0.0;
  double y1[NS];
  step_store<NS>(m, cx, t0, h, y, y1, Kl, (lds_double *)KBl);
  rodas5_step_vjp<NS>(m, cx, t0, h, y, yb, thb, tb, hb, Kl, KBl);
  double *yo = H.ybar + (long long)item * P.n;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    if (rd_type(m.rd[s]) != R_NONE) yo[rd_idx(m.rd[s])] = yb[s];
  if (lane == 0) {
    H.t0bar[item] = tb - hb;
    H.t1bar[item] = hb;
  }
#pragma unroll
  for (int j = 0; j < MAXDIR; ++j)
    if (lane == j && j < H.ndir) H.grad[(long long)item * H.ndir + j] = thb[j];
}

// ---------------------------------------------------------------------------
// launch order of the reverse launch.  Each mode's cost is proportional to its
// accepted steps, which the tape records; handing the longest modes out first
// (longest-processing-time order) shortens the launch's tail, where the last
// long modes run on a few SIMDs while the rest of the GPU idles.  A bucket
// sort over 8-step buckets (the last one open-ended): histogram, exclusive
// scan from the longest bucket down, scatter.  Ties within a bucket take any
// order; each mode's result does not depend on it.
constexpr int ORDER_BUCKETS = 8192;

__device__ __forceinline__ int order_bucket(int32_t count) {
  int b = count < 0 ? 0 : (count >> 3);
  b = b > ORDER_BUCKETS - 1 ? ORDER_BUCKETS - 1 : b;
  return ORDER_BUCKETS - 1 - b;  // bucket 0 holds the longest modes
}

__global__ void __launch_bounds__(256) deb_order_hist_kernel(const int32_t *count, long long n, int32_t *hist) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) atomicAdd(&hist[order_bucket(count[i])], 1);
}

// exclusive prefix sums of the histogram, in place (one block of 1024 threads)
__global__ void __launch_bounds__(1024) deb_order_scan_kernel(int32_t *hist) {
  constexpr int PER = ORDER_BUCKETS / 1024;
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  int32_t loc[PER], sum = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    loc[j] = hist[t * PER + j];
    sum += loc[j];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = part[t] - sum;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    hist[t * PER + j] = run;
    run += loc[j];
  }
}

__global__ void __launch_bounds__(256) deb_order_scatter_kernel(const int32_t *count, long long n, int32_t *offs,
                                                                 int32_t *order) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) order[atomicAdd(&offs[order_bucket(count[i])], 1)] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// host side

// rings of the reverse launch: twice the waves the device can hold at the
// kernel's occupancy
template <int NS>
int adjoint_rings(int &n_rings) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (int r = hip_check(hipGetDevice(&dev), "hipGetDevice")) return r;
  if (int r = hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "CU count"))
    return r;
  if (int r = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &per_cu, reinterpret_cast<const void *>(deb_adjoint_kernel<NS>), WAVE, 0),
                        "occupancy query"))
    return r;
  n_rings = 2 * std::max(per_cu, 1) * std::max(ncu, 1);
  return 0;
}

// workspace: [ring busy flags][order histogram][launch order][rings], each
// part 256-byte aligned
inline long long round256(long long b) { return (b + 255) / 256 * 256; }
inline long long adjoint_hist_off(int n_rings) { return round256((long long)n_rings * 4); }
inline long long adjoint_order_off(int n_rings) { return adjoint_hist_off(n_rings) + round256(4LL * ORDER_BUCKETS); }
inline long long adjoint_rings_off(int n_rings, long long n_modes) {
  return adjoint_order_off(n_rings) + round256(4 * n_modes);
}
inline long long adjoint_ws_bytes(int ns, int ckpt, int n_rings, long long n_modes) {
  return adjoint_rings_off(n_rings, n_modes) + (long long)n_rings * ring_doubles(ns, ckpt) * 8;
}

template <int NS>
int adjoint_workspace(int ckpt, long long n_modes, size_t *bytes) {
  int n_rings = 0;
  if (int r = adjoint_rings<NS>(n_rings)) return r;
  *bytes = (size_t)adjoint_ws_bytes(NS, ckpt, n_rings, n_modes);
  return 0;
}

template <int NS>
int run_adjoint(const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                int64_t stride, const int32_t *desc, const double *q, const double *w, int nk, const double *kmodes,
                int nout, const double *tau_out, int ndir, const double *dcos, const double *dtab,
                const double *dtau, const double *ybar, const int32_t *status, double *grad, int ckpt,
                const void *tape, size_t tape_bytes, void *ws, size_t ws_bytes, hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, n_cosmo, cosmo, tables, stride, desc, q, w, nk, kmodes, nout, tau_out))
    return r;
  AdjArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ndir = ndir;
  A.dcos = dcos;
  A.dtab = dtab;
  A.dtau = dtau;
  A.ybar = ybar;
  A.status = status;
  A.grad = grad;
  if (int r = tape_view(A.tape, const_cast<void *>(tape), tape_bytes, (long long)n_cosmo * nk, NS, ckpt)) return r;
  int n_rings = 0;
  if (int r = adjoint_rings<NS>(n_rings)) return r;
  const long long n_modes = (long long)n_cosmo * nk;
  const long long need = adjoint_ws_bytes(NS, ckpt, n_rings, n_modes);
  if (!ws || (long long)ws_bytes < need) return fail(DEB_E_ARG, "workspace smaller than deb_adjoint_workspace_bytes");
  char *wb = static_cast<char *>(ws);
  A.ring_busy = reinterpret_cast<int32_t *>(wb);
  int32_t *hist = reinterpret_cast<int32_t *>(wb + adjoint_hist_off(n_rings));
  int32_t *order = reinterpret_cast<int32_t *>(wb + adjoint_order_off(n_rings));
  A.order = order;
  A.rings = reinterpret_cast<double *>(wb + adjoint_rings_off(n_rings, n_modes));
  A.n_rings = n_rings;
  A.n_in = output_inputs(*cfg, A.in_idx);
  // (flags and histogram cleared by kernels: no memset nodes if the caller
  // captures the launches in a graph)
  hipLaunchKernelGGL(deb_ring_reset_kernel, dim3((n_rings + 255) / 256), dim3(256), 0, st, A.ring_busy, n_rings);
  if (int r = hip_check(hipGetLastError(), "ring reset launch")) return r;
  hipLaunchKernelGGL(deb_ring_reset_kernel, dim3(ORDER_BUCKETS / 256), dim3(256), 0, st, hist, ORDER_BUCKETS);
  if (int r = hip_check(hipGetLastError(), "histogram reset launch")) return r;
  const unsigned nb = (unsigned)((n_modes + 255) / 256);
  hipLaunchKernelGGL(deb_order_hist_kernel, dim3(nb), dim3(256), 0, st, A.tape.count, n_modes, hist);
  if (int r = hip_check(hipGetLastError(), "order histogram launch")) return r;
  hipLaunchKernelGGL(deb_order_scan_kernel, dim3(1), dim3(1024), 0, st, hist);
  if (int r = hip_check(hipGetLastError(), "order scan launch")) return r;
  hipLaunchKernelGGL(deb_order_scatter_kernel, dim3(nb), dim3(256), 0, st, A.tape.count, n_modes, hist, order);
  if (int r = hip_check(hipGetLastError(), "order scatter launch")) return r;
  hipLaunchKernelGGL(deb_adjoint_kernel<NS>, dim3((unsigned)n_modes), dim3(WAVE), 0, st, P, A);
  return hip_check(hipGetLastError(), "adjoint kernel launch");
}

template <int NS>
int run_vjp_hook(bool step, const deb_config_t *cfg, const Layout &L, const deb_cosmo_t *cosmo, const double *tables,
                 int64_t stride, const int32_t *desc, const double *q, const double *w, const VjpHook &H,
                 hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, 1, cosmo, tables, stride, desc, q, w, 1, nullptr, 1, nullptr)) return r;
  if (step) hipLaunchKernelGGL(deb_step_vjp_kernel<NS>, dim3(H.m), dim3(WAVE), 0, st, P, H);
  else hipLaunchKernelGGL(deb_rhs_vjp_kernel<NS>, dim3(H.m), dim3(WAVE), 0, st, P, H);
  return hip_check(hipGetLastError(), step ? "step vjp kernel launch" : "rhs vjp kernel launch");
}

inline int check_dirs(int ndir, const double *dcos, const double *dtab) {
  if (ndir < 1 || ndir > DEB_MAXDIR) return fail(DEB_E_ARG, "ndir must be in [1, DEB_MAXDIR]");
  if (!dcos || !dtab) return fail(DEB_E_ARG, "null tangent arrays");
  return 0;
}

}  // namespace deb

using namespace deb;

extern "C" {

int deb_tape_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int64_t n_states, int ckpt_every, size_t *bytes) {
  if (!cfg || !bytes || n_cosmo < 1 || nk < 1 || n_states < 0) return fail(DEB_E_ARG, "invalid argument");
  if (ckpt_every < 1 || ckpt_every > TAPE_MAX_CKPT) return fail(DEB_E_ARG, "checkpoint interval outside [1, 4096]");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  const long long n_modes = (long long)n_cosmo * nk;
  // one chunk per started segment of every mode: at most n_states / C + n_modes
  const long long chunks = n_states / ckpt_every + n_modes + 1;
  *bytes = (size_t)(tape_pool_off(n_modes) + chunks * tape_chunk_bytes(L.ns, ckpt_every));
  return 0;
}

int deb_adjoint_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int ckpt_every, size_t *bytes) {
  if (!cfg || !bytes || n_cosmo < 1 || nk < 1) return fail(DEB_E_ARG, "invalid argument");
  if ((long long)n_cosmo * nk > 0x7fffffffLL) return fail(DEB_E_ARG, "more than 2^31 - 1 modes");
  if (ckpt_every < 1 || ckpt_every > TAPE_MAX_CKPT) return fail(DEB_E_ARG, "checkpoint interval outside [1, 4096]");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  DEB_DISPATCH(L.ns, (adjoint_workspace<NS>(ckpt_every, (long long)n_cosmo * nk, bytes)))
}

int deb_solve_adjoint(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                      int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                      const double *kmodes, int nout, const double *tau_out, int ndir, const double *dcosmo,
                      const double *dtables, const double *dtau_out, const double *ybar, const int32_t *status,
                      double *grad, int ckpt_every, const void *tape, size_t tape_bytes, void *workspace,
                      size_t workspace_bytes, void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (int r = check_dirs(ndir, dcosmo, dtables)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !dtau_out || !ybar || !status || !grad)
    return fail(DEB_E_ARG, "invalid mode arrays");
  if (cfg->return_full) return fail(DEB_E_ARG, "the gradient path takes cotangents of the 20 output fields");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_adjoint<NS>(cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes, nout,
                                      tau_out, ndir, dcosmo, dtables, dtau_out, ybar, status, grad, ckpt_every, tape,
                                      tape_bytes, workspace, workspace_bytes, st)))
}

int deb_rhs_vjp(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
                const int32_t *desc, const double *q, const double *w, int ndir, const double *dcosmo,
                const double *dtables, int m, const int32_t *ci, const double *kk, const double *tau, const double *y,
                const double *fbar, double *ybar, double *taubar, double *grad, void *stream) {
  if (int r = check_common(cfg, 1, cosmo, tables, desc, q, w)) return r;
  if (int r = check_dirs(ndir, dcosmo, dtables)) return r;
  if (m < 1 || !ci || !kk || !tau || !y || !fbar || !ybar || !taubar || !grad)
    return fail(DEB_E_ARG, "invalid hook arrays");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  VjpHook H{ci, kk, tau, nullptr, y, fbar, ybar, taubar, nullptr, grad, m, ndir, dcosmo, dtables};
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_vjp_hook<NS>(false, cfg, L, cosmo, tables, table_stride, desc, q, w, H, st)))
}

int deb_rodas5_step_vjp(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables,
                        int64_t table_stride, const int32_t *desc, const double *q, const double *w, int ndir,
                        const double *dcosmo, const double *dtables, int m, const int32_t *ci, const double *kk,
                        const double *t0, const double *t1, const double *y0, const double *y1bar, double *y0bar,
                        double *t0bar, double *t1bar, double *grad, void *stream) {
  if (int r = check_common(cfg, 1, cosmo, tables, desc, q, w)) return r;
  if (int r = check_dirs(ndir, dcosmo, dtables)) return r;
  if (m < 1 || !ci || !kk || !t0 || !t1 || !y0 || !y1bar || !y0bar || !t0bar || !t1bar || !grad)
    return fail(DEB_E_ARG, "invalid hook arrays");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  VjpHook H{ci, kk, t0, t1, y0, y1bar, y0bar, t0bar, t1bar, grad, m, ndir, dcosmo, dtables};
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_vjp_hook<NS>(true, cfg, L, cosmo, tables, table_stride, desc, q, w, H, st)))
}

}  // extern "C"
