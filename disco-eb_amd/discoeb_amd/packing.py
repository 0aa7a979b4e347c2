"""Flatten the reference ``param`` dict into the C-ABI input layout.

``deb_cosmo_t`` (24 doubles per cosmology) and the spline table block
``tables[n_cosmo][stride]`` with its descriptor ``desc[DEB_NSPLINE][2]``
(see include/discoeb_hip.h).  The field order of ``deb_cosmo_t`` and the
spline order are fixed by the header; the meaning of each entry is the
reference's ``param`` key of the same name.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .background import get_neutrino_momentum_bins

COSMO_FIELDS = ("grhom", "grhog", "grhor", "Omegam", "Omegab", "OmegaDE", "Omegak", "Neff", "Nmnu", "amnu",
                "w_DE_0", "w_DE_a", "cs2_DE", "YHe", "H0", "taumin", "A_s", "n_s", "k_p")
N_COSMO_DOUBLES = 24
SPLINE_KEYS = ("cs2a_of_loga_spline", "xe_of_loga_spline", "logrhonu_of_loga_spline", "logpnu_of_loga_spline",
               "tau_of_a_spline", "a_of_tau_spline", "xe_of_tau_spline")
N_OUT_FIELDS = 20

STATUS_OK, STATUS_MAX_STEPS, STATUS_NONFINITE, STATUS_TAPE_FULL, STATUS_NOT_RUN = 0, 1, 2, 3, -1
SOLVER_RODAS5 = 0


class DebConfig(ctypes.Structure):
    _fields_ = [("lmaxg", ctypes.c_int32), ("lmaxgp", ctypes.c_int32), ("lmaxr", ctypes.c_int32),
                ("lmaxnu", ctypes.c_int32), ("nqmax", ctypes.c_int32), ("max_steps", ctypes.c_int32),
                ("return_full", ctypes.c_int32), ("solver", ctypes.c_int32),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("pcoeff", ctypes.c_double),
                ("icoeff", ctypes.c_double), ("dcoeff", ctypes.c_double), ("factormax", ctypes.c_double),
                ("factormin", ctypes.c_double)]


def make_config(*, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4, pcoeff=0.25,
                icoeff=0.80, dcoeff=0.0, factormax=20.0, factormin=0.3, max_steps=2048, return_full=False,
                solver=SOLVER_RODAS5) -> DebConfig:
    return DebConfig(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), int(max_steps),
                     int(bool(return_full)), int(solver), float(rtol), float(atol), float(pcoeff), float(icoeff),
                     float(dcoeff), float(factormax), float(factormin))


def nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax) -> int:
    """Number of state variables (``perturbations.py:739``)."""
    return 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nqmax * (lmaxnu + 1) + 2


def _real(v):
    return np.real(np.asarray(v))


def pack_cosmo(param: dict, part=np.real) -> np.ndarray:
    """(B, 24) float64 deb_cosmo_t records.  ``part`` picks real/imag parts."""
    B = np.asarray(param["H0"]).shape[0]
    out = np.zeros((B, N_COSMO_DOUBLES), dtype=np.float64)
    for j, key in enumerate(COSMO_FIELDS):
        if key == "k_p" and key not in param:
            out[:, j] = 0.05 if part is np.real else 0.0
            continue
        out[:, j] = part(np.broadcast_to(np.asarray(param[key]), (B,)))
    return out


def table_layout(param: dict):
    """Descriptor (DEB_NSPLINE, 2) int32 (offset, n) and the per-cosmology stride."""
    desc = np.zeros((len(SPLINE_KEYS), 2), dtype=np.int32)
    off = 0
    for s, key in enumerate(SPLINE_KEYS):
        n = param[key].n
        desc[s] = (off, n)
        off += 3 * n
    return desc, off


def pack_tables(param: dict, part=np.real):
    """(B, stride) float64 knots (x, y, S per spline) and the descriptor."""
    desc, stride = table_layout(param)
    B = np.asarray(param["H0"]).shape[0]
    tab = np.zeros((B, stride), dtype=np.float64)
    for s, key in enumerate(SPLINE_KEYS):
        sp = param[key]
        off, n = desc[s]
        for j, arr in enumerate((sp.x, sp.y, sp.S)):
            tab[:, off + j * n: off + (j + 1) * n] = part(np.broadcast_to(arr, (B, n)))
    return tab, desc


def nu_nodes(nqmax: int):
    q, w = get_neutrino_momentum_bins(nqmax)
    return np.ascontiguousarray(q, dtype=np.float64), np.ascontiguousarray(w, dtype=np.float64)


def tau_of_aexp(param: dict, aexp_out) -> np.ndarray:
    """tau_out[B, nout] = tau_of_a_spline(aexp_out) (``perturbations.py:975``)."""
    a = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    B = np.asarray(param["H0"]).shape[0]
    return param["tau_of_a_spline"].evaluate(np.broadcast_to(a, (B, a.size)))
