"""Host mirror of DISCO-EB's perturbation driver on the gfx950 kernels.

Same keyword-only signatures, defaults and return arity as the reference:

* ``evolve_perturbations``          src/discoeb/perturbations.py:926-997
* ``evolve_perturbations_batched``  src/discoeb/perturbations.py:1000-1061
* ``get_power``                     src/discoeb/perturbations.py:1101-1123

plus ``evolve_perturbations_multi``: many cosmologies x k-modes in one
launch (one wavefront per mode, largest k first).  Results are torch FP64 tensors on the GPU.  There is no
CPU fallback: a missing GPU or library raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import packing
from .packing import N_OUT_FIELDS, STATUS_OK


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.DebError("the perturbation hot path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class DeviceInputs:
    """Packed inputs of n_cosmo cosmologies resident in device memory."""

    def __init__(self, param: dict, cfg: packing.DebConfig, device=None):
        torch = _torch()
        self.device = torch.device(device if device is not None else "cuda")
        cos = packing.pack_cosmo(param)
        tab, desc = packing.pack_tables(param)
        self.n_cosmo = cos.shape[0]
        self.cosmo = torch.from_numpy(np.ascontiguousarray(cos)).to(self.device)
        self.tables = torch.from_numpy(np.ascontiguousarray(tab)).to(self.device)
        self.stride = tab.shape[1]
        self.desc = np.ascontiguousarray(desc, dtype=np.int32)
        q, w = packing.nu_nodes(cfg.nqmax)
        self.q = np.ascontiguousarray(q if q.size else np.zeros(1))
        self.w = np.ascontiguousarray(w if w.size else np.zeros(1))
        self.cfg = cfg


def solve_forward(inp: DeviceInputs, kmodes, tau_out, *, throw: bool = True, stream=None, batch_size: int = 0):
    """Run deb_solve_forward; kmodes (n_cosmo, nk), tau_out (n_cosmo, nout) tensors.
    With batch_size > 0, deb_solve_forward_batched: groups of batch_size
    consecutive k share one step sequence (evolve_modes_batched)."""
    torch = _torch()
    L = _lib.lib()
    cfg = inp.cfg
    kmodes = kmodes.to(inp.device, torch.float64).contiguous()
    tau_out = tau_out.to(inp.device, torch.float64).contiguous()
    n_cosmo, nk = kmodes.shape
    nout = tau_out.shape[1]
    if n_cosmo != inp.n_cosmo or tau_out.shape[0] != n_cosmo:
        raise ValueError("kmodes/tau_out must have one row per cosmology")
    n = L.deb_nvar(ctypes.byref(cfg))
    width = n if cfg.return_full else N_OUT_FIELDS
    y = torch.empty((n_cosmo, nk, nout, width), dtype=torch.float64, device=inp.device)
    status = torch.full((n_cosmo, nk), -1, dtype=torch.int32, device=inp.device)
    nsteps = torch.zeros((n_cosmo, nk), dtype=torch.int32, device=inp.device)
    wsb = ctypes.c_size_t(0)
    st = stream if stream is not None else torch.cuda.current_stream(inp.device)
    common = (ctypes.byref(cfg), n_cosmo, _lib.ptr(inp.cosmo), _lib.ptr(inp.tables), inp.stride, _lib.ptr(inp.desc),
              _lib.ptr(inp.q), _lib.ptr(inp.w), nk, _lib.ptr(kmodes), nout, _lib.ptr(tau_out))
    if batch_size > 0:
        _lib.check(L.deb_batched_workspace_bytes(ctypes.byref(cfg), n_cosmo, nk, ctypes.byref(wsb)),
                   "deb_batched_workspace_bytes")
        ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=inp.device)
        _lib.check(L.deb_solve_forward_batched(*common, batch_size, _lib.ptr(y), _lib.ptr(status), _lib.ptr(nsteps),
                                               _lib.ptr(ws), wsb.value, ctypes.c_void_p(st.cuda_stream)),
                   "deb_solve_forward_batched")
    else:
        _lib.check(L.deb_workspace_bytes(ctypes.byref(cfg), n_cosmo, nk, nout, ctypes.byref(wsb)),
                   "deb_workspace_bytes")
        ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=inp.device)
        _lib.check(L.deb_solve_forward(*common, _lib.ptr(y), _lib.ptr(status), _lib.ptr(nsteps), _lib.ptr(ws),
                                       wsb.value, ctypes.c_void_p(st.cuda_stream)),
                   "deb_solve_forward")
    if throw:
        st.synchronize()
        bad = (status != STATUS_OK)
        if bool(bad.any()):
            nbad = int(bad.sum())
            codes = sorted(set(status[bad].cpu().tolist()))
            raise _lib.DebError(f"{nbad} mode(s) failed (status codes {codes}: 1 = max_steps reached, "
                                "2 = non-finite error norm); pass throw=False to inspect them")
    return y, status, nsteps


def _kgrid(kmin, kmax, num_k, dologk):
    return np.geomspace(kmin, kmax, num_k) if dologk else np.linspace(kmin, kmax, num_k)


def _config(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin,
            max_steps, return_full):
    return packing.make_config(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, rtol=rtol,
                               atol=atol, pcoeff=pcoeff, icoeff=icoeff, dcoeff=dcoeff, factormax=factormax,
                               factormin=factormin, max_steps=max_steps, return_full=return_full)


def evolve_perturbations(*, param, aexp_out, kmin: float, kmax: float, num_k: int, lmaxg: int = 11,
                         lmaxgp: int = 11, lmaxr: int = 11, lmaxnu: int = 8, nqmax: int = 3, rtol: float = 1e-4,
                         atol: float = 1e-4, pcoeff: float = 0.25, icoeff: float = 0.80, dcoeff: float = 0.0,
                         factormax: float = 20.0, factormin: float = 0.3, max_steps: int = 2048,
                         return_full: bool = False, dologk: bool = True, throw: bool = True):
    """Evolve synchronous-gauge perturbations of one cosmology (``param`` from
    :func:`discoeb_amd.background.evolve_background`).

    Returns ``(y, kmodes, param)`` with ``y`` of shape (num_k, nout, 20) (or
    (num_k, nout, nvar) with ``return_full``) as a torch FP64 tensor on the GPU.
    ``param`` gains the keys lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, nout, tau_out.
    ``throw`` mirrors diffrax: any mode that hits max_steps raises.
    """
    torch = _torch()
    cfg = _config(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin,
                  max_steps, return_full)
    if np.asarray(param["H0"]).size != 1:
        raise ValueError("evolve_perturbations takes one cosmology; use evolve_perturbations_multi")
    kmodes = _kgrid(kmin, kmax, num_k, dologk)
    aexp_out = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    tau_out = packing.tau_of_aexp(param, aexp_out).real
    inp = DeviceInputs(param, cfg)
    y, _, _ = solve_forward(inp, torch.from_numpy(kmodes[None, :]), torch.from_numpy(tau_out), throw=throw)
    param["lmaxg"] = lmaxg
    param["lmaxgp"] = lmaxgp
    param["lmaxr"] = lmaxr
    param["lmaxnu"] = lmaxnu
    param["nqmax"] = nqmax
    param["nout"] = aexp_out.size
    param["tau_out"] = tau_out[0]
    return y[0], torch.from_numpy(kmodes).to(y.device), param


def evolve_perturbations_batched(*, param, aexp_out, kmin: float, kmax: float, num_k: int, lmaxg: int = 11,
                                 lmaxgp: int = 11, lmaxr: int = 11, lmaxnu: int = 8, nqmax: int = 3,
                                 rtol: float = 1e-4, atol: float = 1e-4, pcoeff: float = 0.25, icoeff: float = 0.80,
                                 dcoeff: float = 0.0, factormax: float = 20.0, factormin: float = 0.3,
                                 max_steps: int = 2048, batch_size: int = 16, throw: bool = True):
    """evolve_perturbations_batched (perturbations.py:1000-1061): geomspace
    k-grid split into groups of batch_size consecutive modes, each group
    integrated with one shared step sequence (evolve_modes_batched :820-922);
    ``num_k % batch_size == 0``; returns ``(y, kmodes)``."""
    torch = _torch()
    if num_k % batch_size != 0:
        raise ValueError("num_k must be a multiple of batch_size")
    if np.asarray(param["H0"]).size != 1:
        raise ValueError("evolve_perturbations_batched takes one cosmology")
    cfg = _config(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin,
                  max_steps, False)
    kmodes = np.geomspace(kmin, kmax, num_k)
    aexp_out = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    tau_out = packing.tau_of_aexp(param, aexp_out).real
    inp = DeviceInputs(param, cfg)
    y, _, _ = solve_forward(inp, torch.from_numpy(kmodes[None, :]), torch.from_numpy(tau_out), throw=throw,
                            batch_size=batch_size)
    param["nout"] = aexp_out.size
    return y[0], torch.from_numpy(kmodes).to(y.device)


def evolve_perturbations_multi(*, param, aexp_out, kmodes, lmaxg: int = 11, lmaxgp: int = 11, lmaxr: int = 11,
                               lmaxnu: int = 8, nqmax: int = 3, rtol: float = 1e-4, atol: float = 1e-4,
                               pcoeff: float = 0.25, icoeff: float = 0.80, dcoeff: float = 0.0,
                               factormax: float = 20.0, factormin: float = 0.3, max_steps: int = 2048,
                               return_full: bool = False, throw: bool = True):
    """All (cosmology, k) modes of a batched ``param`` (B cosmologies) in one
    launch.  ``kmodes`` is (nk,) or (B, nk).  Returns
    ``(y[B, nk, nout, width], status[B, nk], nsteps[B, nk])``."""
    torch = _torch()
    cfg = _config(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin,
                  max_steps, return_full)
    B = np.asarray(param["H0"]).shape[0]
    k = np.asarray(kmodes, dtype=np.float64)
    if k.ndim == 1:
        k = np.broadcast_to(k, (B, k.size))
    tau_out = packing.tau_of_aexp(param, aexp_out).real
    inp = DeviceInputs(param, cfg)
    return solve_forward(inp, torch.from_numpy(np.ascontiguousarray(k)), torch.from_numpy(tau_out), throw=throw)


def get_power(*, k, y, idx: int, param: dict):
    """P(k) = 2 pi^2 A_s (k/k_p)^(n_s-1) k^-3 y[..., idx]^2 (perturbations.py:1101-1123)."""
    A_s = float(np.real(np.asarray(param["A_s"])).ravel()[0])
    n_s = float(np.real(np.asarray(param["n_s"])).ravel()[0])
    k_p = float(np.real(np.asarray(param.get("k_p", 0.05))).ravel()[0])
    return 2 * np.pi**2 * A_s * (k / k_p) ** (n_s - 1) * k ** (-3) * y[..., idx] ** 2
