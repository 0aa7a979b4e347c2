// gfx950 kernels and C ABI of the DISCO-EB perturbation hot path.
//
//   deb_forward_kernel  one wavefront per (cosmology, k) item (largest k
//                       dispatched first), runs
//                       determine_starting_time, adiabatic ICs and the
//                       adaptive Rodas5 loop (diffrax PID semantics), and
//                       writes the output fields at every tau_out.
//   deb_forward_tape_kernel  the same solve, also writing every accepted state
//                       to the checkpoint tape of the gradient path.
//   deb_forward_batched_kernel  one wavefront per (cosmology, group of
//                       batch_size consecutive k): the group shares one step
//                       sequence (evolve_perturbations_batched).
//   deb_rhs_kernel      test hook: model_synchronous for given states.
//   deb_step_kernel     test hook: one Rodas5Transformed step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "deb_forward.h"
#include "deb_host.h"

namespace deb {

thread_local std::string g_err;

// Workspace of deb_solve_forward_batched_noalloc: a file-scope device array
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never passes `scratch` itself to a kernel (in host code the name
// is a stand-in, not a device address): scratch_ptr() returns the device
// address from hipGetSymbolAddress, looked up once per device.
constexpr size_t SCRATCH_BYTES = 64ull << 20;
__device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[MAX_DEVICES];

// the current device's address of `scratch` (nullptr and an error message on failure)
static void *scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    fail(DEB_E_HIP, "no current HIP device for the allocation-free workspace");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fail(DEB_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// evolve_perturbations_batched: one wavefront per (cosmology, group of B
// consecutive k), groups of the largest k first.
template <int NS>
__global__ void __launch_bounds__(WAVE) deb_forward_batched_kernel(const Params<NS> P, const int B, double *ws) {
  __shared__ double lds[fwd_lds_doubles<NS>()];
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  const int ng = P.nk / B;
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * ng) return;
  const int g = ng - 1 - item / P.n_cosmo;
  const int ci = item % P.n_cosmo;
  integrate_batch<NS>(P, ci, g * B, B, lds, (lds_double *)pcr, ws);
}

// The forward solve that also writes every accepted state to the checkpoint
// tape of the gradient path (deb_solve_forward_tape).
template <int NS>
__global__ void __launch_bounds__(WAVE) deb_forward_tape_kernel(const Params<NS> P, const TapeView T) {
  __shared__ double lds[fwd_lds_doubles<NS>()];
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * P.nk) return;
  const int ki = P.nk - 1 - item / P.n_cosmo;
  const int ci = item % P.n_cosmo;
  integrate_mode<NS, true>(P, ci, ki, lds, (lds_double *)pcr, &T);
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
  __shared__ double pcr[pcr_lds_doubles<NS>()];
  const int item = blockIdx.x;
  if (item >= H.m) return;
  const int lane = lane_id();
  Ctx cx;
  make_ctx<NS>(cx, P, H.ci[item], H.kk[item]);
  __shared__ ModeU mu;
  Mode<NS> m;
  m.pcr = (lds_double *)pcr;
  m.u = (lds_mode_u *)&mu;
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

// test hook: the start time of every (cosmology, k) mode
// (determine_starting_time and 0.99 min(min tau_out, tau*))
__global__ void __launch_bounds__(WAVE) deb_start_kernel(const Params<1> P, double *ts) {
  const int item = blockIdx.x;
  if (item >= P.n_cosmo * P.nk) return;
  const int ci = item / P.nk;
  Ctx cx;
  make_ctx<1>(cx, P, ci, P.kmodes[item]);
  const double *tout = P.tau_out + (int64_t)ci * P.nout;
  double tmin = tout[0];
  for (int i = 1; i < P.nout; ++i) tmin = fmin(tmin, tout[i]);
  const double t = start_time(cx, tmin);
  if (lane_id() == 0) ts[item] = t;
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
int run_forward_batched(const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo,
                        const double *tables, int64_t stride, const int32_t *desc, const double *q, const double *w,
                        int nk, const double *kmodes, int nout, const double *tau_out, int batch, double *y_out,
                        int32_t *status, int32_t *nsteps, void *ws, hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, n_cosmo, cosmo, tables, stride, desc, q, w, nk, kmodes, nout, tau_out))
    return r;
  P.y_out = y_out;
  P.status = status;
  P.nsteps = nsteps;
  const unsigned items = (unsigned)((int64_t)n_cosmo * (nk / batch));
  hipLaunchKernelGGL(deb_forward_batched_kernel<NS>, dim3(items), dim3(WAVE), 0, st, P, batch,
                     static_cast<double *>(ws));
  return hip_check(hipGetLastError(), "batched forward kernel launch");
}

// clears the tape's chunk counter (a kernel rather than a memset, so that a
// caller may capture the launch sequence in a graph)
__global__ void deb_tape_reset_kernel(unsigned long long *counter) {
  if (threadIdx.x == 0) *counter = 0ull;
}

template <int NS>
int run_forward_tape(const deb_config_t *cfg, const Layout &L, int n_cosmo, const deb_cosmo_t *cosmo,
                     const double *tables, int64_t stride, const int32_t *desc, const double *q, const double *w,
                     int nk, const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                     int32_t *nsteps, int ckpt, void *tape, size_t tape_bytes, hipStream_t st) {
  Params<NS> P;
  if (int r = fill_params<NS>(P, cfg, L, n_cosmo, cosmo, tables, stride, desc, q, w, nk, kmodes, nout, tau_out))
    return r;
  P.y_out = y_out;
  P.status = status;
  P.nsteps = nsteps;
  TapeView tv;
  if (int r = tape_view(tv, tape, tape_bytes, (long long)n_cosmo * nk, NS, ckpt)) return r;
  hipLaunchKernelGGL(deb_tape_reset_kernel, dim3(1), dim3(WAVE), 0, st, tv.counter);
  if (int r = hip_check(hipGetLastError(), "tape reset launch")) return r;
  hipLaunchKernelGGL(deb_forward_tape_kernel<NS>, dim3((unsigned)((int64_t)n_cosmo * nk)), dim3(WAVE), 0, st, P, tv);
  return hip_check(hipGetLastError(), "taped forward kernel launch");
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

}  // namespace deb

using namespace deb;

extern "C" {

int deb_abi_version(void) { return DEB_ABI_VERSION; }

const char *deb_last_error(void) { return g_err.c_str(); }

int deb_nvar(const deb_config_t *cfg) {
  if (!cfg) return fail(DEB_E_ARG, "null config");
  return 12 + cfg->lmaxg + cfg->lmaxgp + cfg->lmaxr + cfg->nqmax * (cfg->lmaxnu + 1);
}

int deb_slots(const deb_config_t *cfg) {
  if (!cfg) return fail(DEB_E_ARG, "null config");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  return L.ns;
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

int deb_batched_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, size_t *bytes) {
  if (!cfg || !bytes || n_cosmo < 1 || nk < 1) return fail(DEB_E_ARG, "invalid argument");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  *bytes = (size_t)n_cosmo * (size_t)nk * 2 * (size_t)L.ns * WAVE * sizeof(double);
  return 0;
}

int deb_solve_forward_batched(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                              int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                              const double *kmodes, int nout, const double *tau_out, int batch_size, double *y_out,
                              int32_t *status, int32_t *nsteps, void *workspace, size_t workspace_bytes,
                              void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !y_out || !status) return fail(DEB_E_ARG, "invalid mode arrays");
  if (batch_size < 1 || nk % batch_size != 0) return fail(DEB_E_ARG, "num_k must be a multiple of batch_size");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  size_t need = 0;
  if (int r = deb_batched_workspace_bytes(cfg, n_cosmo, nk, &need)) return r;
  if (!workspace || workspace_bytes < need) return fail(DEB_E_ARG, "workspace smaller than deb_batched_workspace_bytes");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_forward_batched<NS>(cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes,
                                              nout, tau_out, batch_size, y_out, status, nsteps, workspace, st)))
}

int deb_solve_forward_batched_noalloc(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo,
                                      const double *tables, int64_t table_stride, const int32_t *desc,
                                      const double *q, const double *w, int nk, const double *kmodes, int nout,
                                      const double *tau_out, int batch_size, double *y_out, int32_t *status,
                                      int32_t *nsteps, void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !y_out || !status) return fail(DEB_E_ARG, "invalid mode arrays");
  if (batch_size < 1 || nk % batch_size != 0) return fail(DEB_E_ARG, "num_k must be a multiple of batch_size");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  size_t need = 0;
  if (int r = deb_batched_workspace_bytes(cfg, n_cosmo, nk, &need)) return r;
  if (need > SCRATCH_BYTES)
    return fail(DEB_E_ARG, "n_cosmo x num_k needs " + std::to_string(need) +
                               " workspace bytes; the allocation-free entry point holds " +
                               std::to_string(SCRATCH_BYTES) + " (use deb_solve_forward_batched)");
  void *ws = scratch_ptr();
  if (!ws) return DEB_E_HIP;
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_forward_batched<NS>(cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes,
                                              nout, tau_out, batch_size, y_out, status, nsteps, ws, st)))
}

int deb_solve_forward_tape(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                           int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                           const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                           int32_t *nsteps, int ckpt_every, void *tape, size_t tape_bytes, void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !y_out || !status) return fail(DEB_E_ARG, "invalid mode arrays");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  hipStream_t st = (hipStream_t)stream;
  DEB_DISPATCH(L.ns, (run_forward_tape<NS>(cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes, nout,
                                           tau_out, y_out, status, nsteps, ckpt_every, tape, tape_bytes, st)))
}

int deb_start_times(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                    int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                    const double *kmodes, int nout, const double *tau_out, double *tau_start, void *stream) {
  if (int r = check_common(cfg, n_cosmo, cosmo, tables, desc, q, w)) return r;
  if (nk < 1 || nout < 1 || !kmodes || !tau_out || !tau_start) return fail(DEB_E_ARG, "invalid mode arrays");
  if ((int64_t)n_cosmo * nk > 0x7fffffff) return fail(DEB_E_ARG, "too many modes");
  Layout L;
  if (int r = build_layout(*cfg, L)) return r;
  Params<1> P;
  if (int r = fill_params<1>(P, cfg, L, n_cosmo, cosmo, tables, table_stride, desc, q, w, nk, kmodes, nout, tau_out))
    return r;
  hipLaunchKernelGGL(deb_start_kernel, dim3((unsigned)((int64_t)n_cosmo * nk)), dim3(WAVE), 0, (hipStream_t)stream,
                     P, tau_start);
  return hip_check(hipGetLastError(), "start-time kernel launch");
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
