/*
 * discoeb_hip.h — C ABI of the MI355X (gfx950) DISCO-EB perturbation hot path.
 *
 * The boundary replaces the reference's Python/JAX entry points
 *   evolve_perturbations          src/discoeb/perturbations.py:926-997
 *   evolve_perturbations_batched  src/discoeb/perturbations.py:1000-1061
 *   evolve_one_mode               src/discoeb/perturbations.py:728-781
 *   get_power                     src/discoeb/perturbations.py:1101-1123
 * and, at the operator level, the stepper seam
 *   Rodas5Transformed.step        src/discoeb/ode_integrators_stiff.py:718-834
 *   model_synchronous (the vector field)  src/discoeb/perturbations.py:84-371
 * The reference has no native FFI: its only seam is the Python call above.  The
 * Python host package (disco-eb_amd/discoeb_amd) binds these functions with
 * ctypes; INTEGRATION.md shows the binding.
 *
 * Conventions: plain pointers and sizes, row-major, FP64.  Pointers marked
 * "device" are HIP device (or managed) memory owned by the caller; pointers
 * marked "host" are read during the call only.  The library allocates no
 * device memory and holds none between calls; scratch comes from the caller's
 * workspace.  All work is enqueued on the caller's stream (hipStream_t passed
 * as void*); no call synchronises the stream.  Every function returns 0 on
 * success and a negative DEB_E_* code on invalid arguments or a HIP error;
 * deb_last_error() gives a message (thread-local).
 */
#ifndef DISCOEB_HIP_H
#define DISCOEB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEB_ABI_VERSION 3
#define DEB_MAXNQ 16    /* maximum massive-neutrino momentum bins */
#define DEB_NOUT_FIELDS 20
#define DEB_MAXDIR 6    /* input tangent directions per gradient launch */
#define DEB_OBS_MAXK 2048 /* k points per cosmology of the observables below */

/* Spline tables read on the path (param dict keys of the reference). */
enum {
  DEB_SPL_CS2A_LOGA = 0,     /* param['cs2a_of_loga_spline']      */
  DEB_SPL_XE_LOGA = 1,       /* param['xe_of_loga_spline']        */
  DEB_SPL_LOGRHONU_LOGA = 2, /* param['logrhonu_of_loga_spline']  */
  DEB_SPL_LOGPNU_LOGA = 3,   /* param['logpnu_of_loga_spline']    */
  DEB_SPL_TAU_A = 4,         /* param['tau_of_a_spline']          */
  DEB_SPL_A_TAU = 5,         /* param['a_of_tau_spline']          */
  DEB_SPL_XE_TAU = 6,        /* param['xe_of_tau_spline']         */
  DEB_NSPLINE = 7
};

/* Per-mode status (the reference raises for the whole call; see throw=).
 * A NaN or inf error norm rejects the step (diffrax counts a NaN error as inf)
 * and shrinks it by factormin.  A mode that fails leaves inf in every output
 * it did not reach (diffrax's unreached SaveAt entries). */
enum {
  DEB_STATUS_OK = 0,
  DEB_STATUS_MAX_STEPS = 1, /* diffrax max_steps reached before t1 */
  DEB_STATUS_NONFINITE = 2, /* a non-finite accepted state, or the step size
                               shrunk to nothing after rejections */
  DEB_STATUS_TAPE_FULL = 3, /* checkpoint tape exhausted (gradient path); the
                               forward results are complete, the tape is not */
  DEB_STATUS_NOT_RUN = -1
};

enum {
  DEB_E_ARG = -1,      /* invalid argument                        */
  DEB_E_HIP = -2,      /* HIP runtime error                       */
  DEB_E_UNSUPPORTED = -3 /* hierarchy sizes outside the kernel's layout */
};

enum { DEB_SOLVER_RODAS5 = 0 };

/* Cosmology scalars (param dict keys of the reference, same meaning). */
typedef struct {
  double grhom, grhog, grhor;
  double Omegam, Omegab, OmegaDE, Omegak;
  double Neff, Nmnu, amnu;
  double w_DE_0, w_DE_a, cs2_DE;
  double YHe, H0, taumin;
  double A_s, n_s, k_p;
  double reserved[5];
} deb_cosmo_t; /* 24 doubles */

/* Hierarchy sizes and adaptive-solver settings (evolve_perturbations kwargs). */
typedef struct {
  int32_t lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax;
  int32_t max_steps;
  int32_t return_full; /* 0: 20 output fields per output time; 1: raw state */
  int32_t solver;      /* DEB_SOLVER_RODAS5 */
  double rtol, atol;
  double pcoeff, icoeff, dcoeff, factormax, factormin;
} deb_config_t;

/* ABI version (DEB_ABI_VERSION). */
int deb_abi_version(void);

/* Message for the last failing call on this host thread. */
const char *deb_last_error(void);

/* Number of state variables n (nvar of evolve_one_mode, perturbations.py:739). */
int deb_nvar(const deb_config_t *cfg);

/* Lane slots per mode (the state occupies deb_slots x 64 doubles per
 * wavefront; a checkpoint-tape chunk holds 16 such states). */
int deb_slots(const deb_config_t *cfg);

/* Bytes of device workspace deb_solve_forward needs (caller allocates). */
int deb_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int nout, size_t *bytes);

/*
 * Forward solve of every (cosmology, k) mode: start time, adiabatic initial
 * conditions, adaptive Rodas5 integration to max(tau_out) with outputs
 * linearly interpolated at tau_out, conversion to the 20 output fields.
 * Replaces evolve_perturbations (perturbations.py:926-997) for n_cosmo == 1
 * and its multi-cosmology generalisation for n_cosmo > 1 (one launch, one
 * wavefront per (cosmology, k) mode, largest k first).
 *
 *   cosmo     device  [n_cosmo] deb_cosmo_t
 *   tables    device  [n_cosmo][table_stride] doubles: for spline s the knots
 *                     x, y, S (second derivatives) start at desc[2s] and each
 *                     have desc[2s+1] entries
 *   desc      host    [DEB_NSPLINE][2] int32 (offset, n) inside one cosmology
 *   q, w      host    [nqmax] massive-nu momentum nodes / kernel weights
 *                     (get_neutrino_momentum_bins, background.py:13-44)
 *   kmodes    device  [n_cosmo][nk]  wavenumbers [1/Mpc]
 *   tau_out   device  [n_cosmo][nout] increasing conformal output times
 *   y_out     device  [n_cosmo][nk][nout][20 or nvar]
 *   status    device  [n_cosmo][nk] int32 DEB_STATUS_*
 *   nsteps    device  [n_cosmo][nk] int32 attempted steps (may be NULL)
 *   workspace device  >= deb_workspace_bytes
 */
int deb_solve_forward(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                      int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                      const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                      int32_t *nsteps, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Replaces evolve_perturbations_batched / evolve_modes_batched
 * (perturbations.py:820-922, 1000-1061): the nk wavenumbers of each cosmology
 * are split into nk / batch_size groups of consecutive k; each group shares one
 * step sequence (start time = the group's minimum start time,
 * determine_start_times_per_batch :786-803; the step-size controller sees the
 * RMS over all the group's filtered, weighted errors,
 * rms_norm_filtered_batched :690-694).  One wavefront per (cosmology, group).
 * Arguments as deb_solve_forward; nk must be a multiple of batch_size
 * (DEB_E_ARG otherwise); status and nsteps are the group's for every mode;
 * workspace >= deb_batched_workspace_bytes (two states per mode).
 */
int deb_batched_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, size_t *bytes);
int deb_solve_forward_batched(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                              int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                              const double *kmodes, int nout, const double *tau_out, int batch_size, double *y_out,
                              int32_t *status, int32_t *nsteps, void *workspace, size_t workspace_bytes,
                              void *stream);

/*
 * deb_solve_forward_batched without a workspace argument: the two states per
 * mode live in a 64 MiB device array inside the library (each GPU holds its
 * own copy until the process ends), so the call allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Limit:
 * n_cosmo * nk * 2 * deb_slots * 512 bytes <= 64 MiB (DEB_E_ARG beyond; e.g.
 * 64 cosmologies x 256 k at lmax 32).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other host threads)
 * race and must be serialised by the caller.  Arguments as
 * deb_solve_forward_batched.
 */
int deb_solve_forward_batched_noalloc(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo,
                                      const double *tables, int64_t table_stride, const int32_t *desc,
                                      const double *q, const double *w, int nk, const double *kmodes, int nout,
                                      const double *tau_out, int batch_size, double *y_out, int32_t *status,
                                      int32_t *nsteps, void *stream);

/*
 * Test hook: the start time of every (cosmology, k) mode, as deb_solve_forward
 * computes it: determine_starting_time (perturbations.py:630-681, bisection
 * util.py:365-396) and tau_start = 0.99 min(min tau_out, tau*) (:743).
 * Arguments as deb_solve_forward; tau_start device [n_cosmo][nk].
 */
int deb_start_times(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                    int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                    const double *kmodes, int nout, const double *tau_out, double *tau_start, void *stream);

/*
 * Operator-level test hook: f = model_synchronous(tau, y) (perturbations.py:84-371)
 * for m independent (cosmology index, k, tau, y) tuples, evaluated by the
 * kernel's wave-parallel right-hand side.
 *   ci [m] int32, kk [m], tau [m], y [m][nvar], f [m][nvar]: device.
 */
int deb_rhs(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
            const int32_t *desc, const double *q, const double *w, int m, const int32_t *ci, const double *kk,
            const double *tau, const double *y, double *f, void *stream);

/*
 * Operator-level test hook: one Rodas5Transformed step (ode_integrators_stiff.py:718-834)
 * from (t0, y0) to t1 for m tuples, using the kernel's structured factorisation.
 *   ci [m], kk [m], t0 [m], t1 [m], y0 [m][nvar], y1 [m][nvar], err [m][nvar]: device.
 */
int deb_rodas5_step(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
                    const int32_t *desc, const double *q, const double *w, int m, const int32_t *ci,
                    const double *kk, const double *t0, const double *t1, const double *y0, double *y1, double *err,
                    void *stream);

/*
 * ---- gradient path ----
 * The reference differentiates evolve_perturbations with jax.jacfwd through
 * diffrax's loop and Rodas5Transformed.step (nb_minimal_example.ipynb:297,
 * nb_discoeb_autodiff.ipynb:274; the solve's DirectAdjoint at
 * perturbations.py:765).  These entry points return the same derivative of
 * the discrete solution in reverse mode, batched over all modes: a taped
 * forward solve, then one backward launch.  The step-size controller's
 * decisions carry no derivative: each accepted step keeps its fraction of
 * [t0, t1] while the start time t0, the last output time t1, the initial
 * state, the cosmology record and every spline knot move.
 *
 * Parameter derivatives are projected on ndir <= DEB_MAXDIR caller-supplied
 * input tangents: for direction j, dcosmo[c][j] (24 doubles, the layout of
 * deb_cosmo_t), dtables[c][j][table_stride] (every knot x, y, S) and
 * dtau_out[c][j][nout].  grad[c][k][j] = sum over the mode's outputs of
 * ybar . d(y_out)/d(direction j).
 */

/*
 * Checkpoint tape: the accepted states of every mode are grouped in segments
 * of ckpt_every consecutive steps; the tape keeps each segment's first state
 * and every step's time (one 512 NS-byte state per segment plus 8 bytes per
 * step, NS = deb_slots), and the reverse launch
 * recomputes each segment forward from its checkpoint.  ckpt_every = 1 keeps
 * every state.
 *
 * Bytes of a tape for n_states accepted states in total (sum over modes of
 * accepted steps + 1).  A sufficient bound: sum of nsteps + n_modes.
 */
int deb_tape_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int64_t n_states, int ckpt_every, size_t *bytes);

/*
 * deb_solve_forward that also records every mode on the checkpoint tape
 * (device memory of tape_bytes, owned by the caller).  Modes whose segments do
 * not fit get DEB_STATUS_TAPE_FULL (their outputs are complete).  Arguments as
 * deb_solve_forward; ckpt_every in [1, 4096]; return_full must be 0 for a
 * later deb_solve_adjoint.
 */
int deb_solve_forward_tape(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                           int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                           const double *kmodes, int nout, const double *tau_out, double *y_out, int32_t *status,
                           int32_t *nsteps, int ckpt_every, void *tape, size_t tape_bytes, void *stream);

/*
 * Bytes of device workspace deb_solve_adjoint needs for n_cosmo x nk modes:
 * one ring per wave the current device can hold at the reverse kernel's
 * occupancy, times two (each ring holds a segment's ckpt_every + 1 states and
 * 8 ckpt_every stage vectors, 512 NS (9 ckpt_every + 1) bytes), plus the
 * launch order of the modes (longest first, from the tape's step counts; 4
 * bytes per mode and a 32 KB histogram).  Queries the current HIP device.
 */
int deb_adjoint_workspace_bytes(const deb_config_t *cfg, int n_cosmo, int nk, int ckpt_every, size_t *bytes);

/*
 * Reverse sweep over a tape written by deb_solve_forward_tape with the same
 * inputs and ckpt_every.  ybar device [n_cosmo][nk][nout][20] cotangents of
 * y_out; status device [n_cosmo][nk] from the taped solve (modes not OK get
 * NaN); grad device [n_cosmo][nk][ndir]; dcosmo, dtables, dtau_out device
 * (see above); workspace device >= deb_adjoint_workspace_bytes (its contents
 * need no initialisation).
 */
int deb_solve_adjoint(const deb_config_t *cfg, int n_cosmo, const deb_cosmo_t *cosmo, const double *tables,
                      int64_t table_stride, const int32_t *desc, const double *q, const double *w, int nk,
                      const double *kmodes, int nout, const double *tau_out, int ndir, const double *dcosmo,
                      const double *dtables, const double *dtau_out, const double *ybar, const int32_t *status,
                      double *grad, int ckpt_every, const void *tape, size_t tape_bytes, void *workspace,
                      size_t workspace_bytes, void *stream);

/*
 * Operator-level test hook: reverse of model_synchronous (perturbations.py:84-371)
 * for m tuples: ybar = (df/dy)^T fbar, taubar = fbar . df/dtau,
 * grad[m][ndir] = fbar . df/d(direction j) at fixed (tau, y).  Tangent arrays
 * are indexed by the tuple's cosmology index ci.  All arrays device.
 */
int deb_rhs_vjp(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables, int64_t table_stride,
                const int32_t *desc, const double *q, const double *w, int ndir, const double *dcosmo,
                const double *dtables, int m, const int32_t *ci, const double *kk, const double *tau, const double *y,
                const double *fbar, double *ybar, double *taubar, double *grad, void *stream);

/*
 * Operator-level test hook: reverse of one Rodas5Transformed step
 * (ode_integrators_stiff.py:718-834) from (t0, y0) to t1: y0bar = (dy1/dy0)^T y1bar,
 * t0bar, t1bar = y1bar . dy1/dt0, dy1/dt1, grad[m][ndir] = y1bar . dy1/d(direction j).
 */
int deb_rodas5_step_vjp(const deb_config_t *cfg, const deb_cosmo_t *cosmo, const double *tables,
                        int64_t table_stride, const int32_t *desc, const double *q, const double *w, int ndir,
                        const double *dcosmo, const double *dtables, int m, const int32_t *ci, const double *kk,
                        const double *t0, const double *t1, const double *y0, const double *y1bar, double *y0bar,
                        double *t0bar, double *t1bar, double *grad, void *stream);

/*
 * Observables of the solution (one output time): y is device
 * [n_cosmo][nk][nfield] (the fields of convert_to_output_variables; field 4 is
 * delta_m, 5 theta_m/(aH)), k device [n_cosmo][nk], A_s, n_s, k_p, bias device
 * [n_cosmo].  The primordial factor is 2 pi^2 A_s (k/k_p)^(n_s-1) k^-3.
 */

/* get_power (perturbations.py:1101-1123): P[c][i] = factor * y[c][i][idx]^2. */
int deb_power(int n_cosmo, int nk, const double *k, const double *y, int nfield, int idx, const double *A_s,
              const double *n_s, const double *k_p, double *P, void *stream);

/* Device scratch (bytes) of the smoothing coefficients: (polyorder + 1) x
 * window_length doubles per cosmology. */
size_t deb_observables_scratch_bytes(int n_cosmo, int window_length, int polyorder);

/* get_power_smoothed (perturbations.py:1126-1159): exp of the Savitzky-Golay
 * filter (util.py:407-444; window_length, polyorder as there, window_length
 * <= nk <= DEB_OBS_MAXK) of log P, with the first window_length/2 and the last
 * (window_length+1)/2 points left unsmoothed.  The caller computes
 * window_length from dlogk as the reference does. */
int deb_power_smoothed(int n_cosmo, int nk, const double *k, const double *y, int nfield, int idx,
                       const double *A_s, const double *n_s, const double *k_p, int window_length, int polyorder,
                       void *scratch, size_t scratch_bytes, double *P_smooth, void *stream);

/* power_Kaiser (perturbations.py:1162-1198): Pkmu[c][i][j] = (bias delta -
 * mu_j^2 theta)^2 for mu device [nmu].  window_length 0: delta, theta from the
 * fields and the primordial factor; otherwise from the smoothed spectra of
 * fields 4 and 5 (delta = +sqrt, theta = -sqrt), as smooth_dlogk does. */
int deb_power_kaiser(int n_cosmo, int nk, int nmu, const double *k, const double *y, int nfield,
                     const double *A_s, const double *n_s, const double *k_p, const double *bias, const double *mu,
                     int window_length, int polyorder, void *scratch, size_t scratch_bytes, double *Pkmu,
                     void *stream);

/* power_multipoles (perturbations.py:1202-1224): P0, P2, P4 [n_cosmo][nk]. */
int deb_power_multipoles(int n_cosmo, int nk, const double *k, const double *y, int nfield, const double *A_s,
                         const double *n_s, const double *k_p, const double *bias, double *P0, double *P2,
                         double *P4, void *stream);

/* get_xi_from_P (perturbations.py:1065-1098): FFTlog of Pk [n_cosmo][N] on the
 * log-spaced k [n_cosmo][N] (N even, <= DEB_OBS_MAXK) for multipole ell;
 * xi and r [n_cosmo][N] in increasing r. */
int deb_xi_from_power(int n_cosmo, int N, const double *k, const double *Pk, int ell, double *xi, double *r,
                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DISCOEB_HIP_H */
